// bb_dispatch.hpp -- run-time degree -> compile-time template dispatch.
#pragma once
#include <type_traits>

#include "bb_common.hpp"

namespace bb {

template <int LO, int HI, typename Fn> int dispatch_n(int N, Fn &&fn) {
    if constexpr (LO > HI) {
        (void)N;
        (void)fn;
        return BB_ERR_ARG;
    } else {
        if (N == LO) return fn(std::integral_constant<int, LO>{});
        return dispatch_n<LO + 1, HI>(N, fn);
    }
}

inline unsigned grid1d(int64_t work, int per_block) { return (unsigned)((work + per_block - 1) / per_block); }

}  // namespace bb

// plain-path helpers shared by the solver (bb_ops.hip)
int bb_pde_fdir3(bb_ctx *ctx, int pde, int N, int64_t K, const double *q, double *fx, double *fy, double *fz);
int bb_plain_combine(bb_ctx *ctx, int pde, int N, int64_t K, double *du, double *const d[9], const double *geo,
                     const double *J);
int bb_lsrk_update(bb_ctx *ctx, int64_t n, double *q, double *res, const double *du, double a, double b,
                   double dt, int first);
