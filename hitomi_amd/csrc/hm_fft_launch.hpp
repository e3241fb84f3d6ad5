// hm_fft_launch.hpp -- host-side dispatch of the LDS Stockham kernels over
// the supported line lengths (compile-time radix plans).
#pragma once

#include "hm_internal.hpp"

#define HM_POW2_LENGTHS(X) X(8) X(16) X(32) X(64) X(128) X(256) X(512) X(1024)
#define HM_MIXED_LENGTHS(X) X(24) X(40) X(48) X(72) X(80) X(96) X(120) X(144) X(160) X(192) X(200) X(216) X(240) X(270) X(288) X(320) X(360) X(384) X(400) X(432) X(480) X(576) X(600) X(640) X(720) X(768) X(800) X(960)

namespace hm {

template <typename T, int L, int DIR, class Src>
int launch_cols_L(hm_ctx* ctx, const char* name, const Src& src, C<T>* dst, const ColGeom& g) {
    constexpr int B0 = ColTile<T, L>::B;
    if (g.n_inner <= 0 || g.n_outer <= 0) return 0;
    if (g.n_outer > 65535) {
        ctx->set_error("fft column pass: outer extent exceeds 65535");
        return -1;
    }
    const C<T>* tw = ctx->twiddle<T>(L);
    if (!tw) return -1;
    if constexpr (has_wide_cols<T, L>()) {
        constexpr int BW = 2 * B0;
        const size_t lds = ((size_t)BW * LdsPad<L>::LP + L) * sizeof(C<T>);  // tile + twiddle table
        if (ctx->wide_cols(L, name) && lds <= ctx->lds_max) {
            dim3 grid((unsigned)((g.n_inner + BW - 1) / BW), (unsigned)g.n_outer, (unsigned)(g.nf > 0 ? g.nf : 1));
            ColGeom gc = g;
            const long long ct = compact_tiles(g, BW);
            if (ct > 0 && ct <= 0x7fffffffLL) {
                gc.compact = 1;
                grid = dim3((unsigned)ct, 1, 1);
            }
            return ctx->launch(name, [&] {
                hipLaunchKernelGGL((k_fft_cols<T, L, DIR, Src, BW>), grid, dim3(FFT_NT), lds, ctx->stream, src, dst, gc,
                                   tw);
            });
        }
    }
    dim3 grid((unsigned)((g.n_inner + B0 - 1) / B0), (unsigned)g.n_outer, (unsigned)(g.nf > 0 ? g.nf : 1));
    ColGeom gc = g;
    const long long ct = compact_tiles(g, B0);
    if (ct > 0 && ct <= 0x7fffffffLL) {
        gc.compact = 1;
        grid = dim3((unsigned)ct, 1, 1);
    }
    const size_t lds = col_lds<T, L>();
    return ctx->launch(name, [&] {
        hipLaunchKernelGGL((k_fft_cols<T, L, DIR, Src>), grid, dim3(FFT_NT), lds, ctx->stream, src, dst, gc, tw);
    });
}

// Lengths without a compile-time plan: the run-time radix plan.
template <typename T, int DIR, class Src>
int launch_cols_rt(hm_ctx* ctx, const char* name, int L, const Src& src, C<T>* dst, const ColGeom& g) {
    RtPlan pl;
    if (!rt_plan_make(L, RtMax<T>::PTS, (int)(128 / sizeof(C<T>)), pl)) {
        ctx->set_error("fft: unsupported line length " + std::to_string(L));
        return -1;
    }
    if (g.n_inner <= 0 || g.n_outer <= 0) return 0;
    if (g.n_outer > 65535) {
        ctx->set_error("fft column pass: outer extent exceeds 65535");
        return -1;
    }
    const C<T>* tw = ctx->twiddle<T>(L);
    if (!tw) return -1;
    dim3 grid((unsigned)((g.n_inner + pl.B - 1) / pl.B), (unsigned)g.n_outer, (unsigned)(g.nf > 0 ? g.nf : 1));
    const size_t lds = (size_t)pl.B * pl.LP * sizeof(C<T>);
    return ctx->launch(name, [&] {
        hipLaunchKernelGGL((k_fft_cols_rt<T, DIR, Src>), grid, dim3(FFT_NT), lds, ctx->stream, src, dst, g, tw, pl);
    });
}

template <typename T, int DIR>
int launch_rows_rt(hm_ctx* ctx, const char* name, int L, const C<T>* src, C<T>* dst, const RowGeom& g) {
    RtPlan pl;
    if (!rt_plan_make(L, RtMax<T>::PTS, 4096, pl)) {
        ctx->set_error("fft: unsupported line length " + std::to_string(L));
        return -1;
    }
    if (g.n_rows <= 0) return 0;
    const C<T>* tw = ctx->twiddle<T>(L);
    if (!tw) return -1;
    dim3 grid((unsigned)((g.n_rows + pl.B - 1) / pl.B));
    const size_t lds = (size_t)pl.B * pl.LP * sizeof(C<T>);
    return ctx->launch(name, [&] {
        hipLaunchKernelGGL((k_fft_rows_rt<T, DIR>), grid, dim3(FFT_NT), lds, ctx->stream, src, dst, g, tw, pl);
    });
}

template <typename T, class RSrc>
int launch_r2c_rt(hm_ctx* ctx, const char* name, int L2, const RSrc& src, C<T>* dst, long long n_rows) {
    RtPlan pl;
    if (!rt_plan_make(L2, RtMax<T>::PTS, 4096, pl)) {
        ctx->set_error("fft r2c: unsupported half length " + std::to_string(L2));
        return -1;
    }
    if (n_rows <= 0) return 0;
    const C<T>* tw = ctx->twiddle<T>(L2);
    const C<T>* tw2 = ctx->twiddle<T>(2 * L2);
    if (!tw || !tw2) return -1;
    dim3 grid((unsigned)((n_rows + pl.B - 1) / pl.B));
    const size_t lds = (size_t)pl.B * pl.LP * sizeof(C<T>);
    return ctx->launch(name, [&] {
        hipLaunchKernelGGL((k_r2c_rows_rt<T, RSrc>), grid, dim3(FFT_NT), lds, ctx->stream, src, dst, n_rows, tw, tw2,
                           pl);
    });
}

template <typename T, class HSrc>
int launch_c2r_rt(hm_ctx* ctx, const char* name, int L2, const HSrc& src, T* dst, long long n_rows, int point_s) {
    RtPlan pl;
    if (!rt_plan_make(L2, RtMax<T>::PTS, 4096, pl)) {
        ctx->set_error("fft c2r: unsupported half length " + std::to_string(L2));
        return -1;
    }
    if (n_rows <= 0) return 0;
    const C<T>* tw = ctx->twiddle<T>(L2);
    const C<T>* tw2 = ctx->twiddle<T>(2 * L2);
    if (!tw || !tw2) return -1;
    dim3 grid((unsigned)((n_rows + pl.B - 1) / pl.B));
    const size_t lds = (size_t)pl.B * pl.LP * sizeof(C<T>);
    return ctx->launch(name, [&] {
        hipLaunchKernelGGL((k_c2r_rows_rt<T, HSrc>), grid, dim3(FFT_NT), lds, ctx->stream, src, dst, n_rows, point_s,
                           tw, tw2, pl);
    });
}

template <typename T, int DIR, class Src>
int launch_cols(hm_ctx* ctx, const char* name, int L, const Src& src, C<T>* dst, const ColGeom& g) {
    switch (L) {
#define HM_CASE(LL) \
    case LL: return launch_cols_L<T, LL, DIR, Src>(ctx, name, src, dst, g);
        HM_POW2_LENGTHS(HM_CASE)
        HM_MIXED_LENGTHS(HM_CASE)
#undef HM_CASE
    default: break;
    }
    return launch_cols_rt<T, DIR, Src>(ctx, name, L, src, dst, g);
}

template <typename T, int L, int DIR>
int launch_rows_L(hm_ctx* ctx, const char* name, const C<T>* src, C<T>* dst, const RowGeom& g) {
    constexpr int B = RowTile<T, L>::B;
    if (g.n_rows <= 0) return 0;
    dim3 grid((unsigned)((g.n_rows + B - 1) / B));
    const size_t lds = row_lds<T, L>();
    const C<T>* tw = ctx->twiddle<T>(L);
    if (!tw) return -1;
    return ctx->launch(name, [&] {
        hipLaunchKernelGGL((k_fft_rows<T, L, DIR>), grid, dim3(FFT_NT), lds, ctx->stream, src, dst, g, tw);
    });
}

template <typename T, int DIR>
int launch_rows(hm_ctx* ctx, const char* name, int L, const C<T>* src, C<T>* dst, const RowGeom& g) {
    switch (L) {
#define HM_CASE(LL) \
    case LL: return launch_rows_L<T, LL, DIR>(ctx, name, src, dst, g);
        HM_POW2_LENGTHS(HM_CASE)
        HM_MIXED_LENGTHS(HM_CASE)
#undef HM_CASE
    default: break;
    }
    return launch_rows_rt<T, DIR>(ctx, name, L, src, dst, g);
}

// r2c rows of N = 2*L2 reals.
template <typename T, int L2, class RSrc>
int launch_r2c_L(hm_ctx* ctx, const char* name, const RSrc& src, C<T>* dst, long long n_rows) {
    constexpr int B = RowTile<T, L2>::B;
    if (n_rows <= 0) return 0;
    dim3 grid((unsigned)((n_rows + B - 1) / B));
    const size_t lds = row_lds<T, L2>();
    const C<T>* tw = ctx->twiddle<T>(L2);
    const C<T>* tw2 = ctx->twiddle<T>(2 * L2);
    if (!tw || !tw2) return -1;
    return ctx->launch(name, [&] {
        hipLaunchKernelGGL((k_r2c_rows<T, L2, RSrc>), grid, dim3(FFT_NT), lds, ctx->stream, src, dst, n_rows, tw, tw2);
    });
}

template <typename T, class RSrc>
int launch_r2c(hm_ctx* ctx, const char* name, int L2, const RSrc& src, C<T>* dst, long long n_rows) {
    switch (L2) {
#define HM_CASE(LL) \
    case LL: return launch_r2c_L<T, LL, RSrc>(ctx, name, src, dst, n_rows);
        HM_POW2_LENGTHS(HM_CASE)
#undef HM_CASE
    default: break;
    }
    return launch_r2c_rt<T, RSrc>(ctx, name, L2, src, dst, n_rows);
}

template <typename T, int L2, class HSrc>
int launch_c2r_L(hm_ctx* ctx, const char* name, const HSrc& src, T* dst, long long n_rows, int point_s) {
    constexpr int B = RowTile<T, L2>::B;
    if (n_rows <= 0) return 0;
    dim3 grid((unsigned)((n_rows + B - 1) / B));
    const size_t lds = row_lds<T, L2>();
    const C<T>* tw = ctx->twiddle<T>(L2);
    const C<T>* tw2 = ctx->twiddle<T>(2 * L2);
    if (!tw || !tw2) return -1;
    return ctx->launch(name, [&] {
        hipLaunchKernelGGL((k_c2r_rows<T, L2, HSrc>), grid, dim3(FFT_NT), lds, ctx->stream, src, dst, n_rows, point_s,
                           tw, tw2);
    });
}

template <typename T, class HSrc>
int launch_c2r(hm_ctx* ctx, const char* name, int L2, const HSrc& src, T* dst, long long n_rows, int point_s) {
    switch (L2) {
#define HM_CASE(LL) \
    case LL: return launch_c2r_L<T, LL, HSrc>(ctx, name, src, dst, n_rows, point_s);
        HM_POW2_LENGTHS(HM_CASE)
#undef HM_CASE
    default: break;
    }
    return launch_c2r_rt<T, HSrc>(ctx, name, L2, src, dst, n_rows, point_s);
}

// Rows from a functor source (forward or backward), any supported length.
template <typename T, int L, int DIR, class RSrc>
int launch_rows_f_L(hm_ctx* ctx, const char* name, const RSrc& src, C<T>* dst, const RowGeom& g) {
    constexpr int B = RowTile<T, L>::B;
    if (g.n_rows <= 0) return 0;
    dim3 grid((unsigned)((g.n_rows + B - 1) / B));
    const size_t lds = row_lds<T, L>();
    const C<T>* tw = ctx->twiddle<T>(L);
    if (!tw) return -1;
    return ctx->launch(name, [&] {
        hipLaunchKernelGGL((k_fft_rows_f<T, L, DIR, RSrc>), grid, dim3(FFT_NT), lds, ctx->stream, src, dst, g, tw);
    });
}

template <typename T, int DIR, class RSrc>
int launch_rows_f(hm_ctx* ctx, const char* name, int L, const RSrc& src, C<T>* dst, const RowGeom& g) {
    switch (L) {
#define HM_CASE(LL) \
    case LL: return launch_rows_f_L<T, LL, DIR, RSrc>(ctx, name, src, dst, g);
        HM_POW2_LENGTHS(HM_CASE)
        HM_MIXED_LENGTHS(HM_CASE)
#undef HM_CASE
    default: break;
    }
    RtPlan pl;
    if (!rt_plan_make(L, RtMax<T>::PTS, 4096, pl)) {
        ctx->set_error("fft: unsupported line length " + std::to_string(L));
        return -1;
    }
    if (g.n_rows <= 0) return 0;
    const C<T>* tw = ctx->twiddle<T>(L);
    if (!tw) return -1;
    dim3 grid((unsigned)((g.n_rows + pl.B - 1) / pl.B));
    const size_t lds = (size_t)pl.B * pl.LP * sizeof(C<T>);
    return ctx->launch(name, [&] {
        hipLaunchKernelGGL((k_fft_rows_f_rt<T, DIR, RSrc>), grid, dim3(FFT_NT), lds, ctx->stream, src, dst, g, tw, pl);
    });
}

}  // namespace hm
