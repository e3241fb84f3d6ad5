// hm_fft.hpp -- LDS-staged Stockham FFT building blocks for gfx950.
//
// A workgroup (256 lanes = 4 wave64) loads a tile of B lines of length L
// into LDS, runs the Stockham autosort stages for a compile-time radix plan
// (radices 2, 3, 4, 5, 6, 8, 9), and stores the tile back.  Two tile shapes:
//   * columns: lines whose points are PS elements apart while B consecutive
//     lines are contiguous in memory, so each point of the tile is one
//     coalesced run of B elements (B*8 = 128 B for FP32, B*16 for FP64);
//   * rows: B contiguous lines.
// Points outside a "support" |n| <= s (n the signed index) are taken as zero
// without being read: this prunes the passes of band-limited inverse
// transforms.
//
// Sign convention is FFTW's (fftw_plan_dft_3d as called at field.hpp:321 and
// field.hpp:339): DIR = -1 forward exp(-2 pi i nk/L), DIR = +1 backward,
// unnormalised.
#pragma once

#include "hm_common.hpp"

#include <type_traits>

namespace hm {

constexpr int FFT_NT = 256;  // lanes per workgroup

// ---------------------------------------------------------------------------
// small DFTs, y_u = sum_t a_t exp(DIR 2 pi i t u / R)
// ---------------------------------------------------------------------------
template <typename T, int R, int DIR> struct Dft;

template <typename T, int DIR> struct Dft<T, 2, DIR> {
    __device__ static inline void run(C<T>* a) {
        C<T> t = a[1];
        a[1] = csub<T>(a[0], t);
        a[0] = cadd<T>(a[0], t);
    }
};

template <typename T, int DIR> struct Dft<T, 4, DIR> {
    __device__ static inline void run(C<T>* a) {
        C<T> t0 = cadd<T>(a[0], a[2]), t1 = csub<T>(a[0], a[2]);
        C<T> t2 = cadd<T>(a[1], a[3]), t3 = cmuli<T>(csub<T>(a[1], a[3]), DIR);
        a[0] = cadd<T>(t0, t2);
        a[2] = csub<T>(t0, t2);
        a[1] = cadd<T>(t1, t3);
        a[3] = csub<T>(t1, t3);
    }
};

template <typename T, int DIR> struct Dft<T, 8, DIR> {
    __device__ static inline void run(C<T>* a) {
        const T h = (T)0.70710678118654752440;
        C<T> e[4] = {a[0], a[2], a[4], a[6]};
        C<T> o[4] = {a[1], a[3], a[5], a[7]};
        Dft<T, 4, DIR>::run(e);
        Dft<T, 4, DIR>::run(o);
        // w^1 = h (1 + DIR i), w^2 = DIR i, w^3 = h (-1 + DIR i)
        C<T> o1 = mk<T>(h * (o[1].x - (T)DIR * o[1].y), h * (o[1].y + (T)DIR * o[1].x));
        C<T> o2 = cmuli<T>(o[2], DIR);
        C<T> o3 = mk<T>(h * (-o[3].x - (T)DIR * o[3].y), h * (-o[3].y + (T)DIR * o[3].x));
        a[0] = cadd<T>(e[0], o[0]);
        a[4] = csub<T>(e[0], o[0]);
        a[1] = cadd<T>(e[1], o1);
        a[5] = csub<T>(e[1], o1);
        a[2] = cadd<T>(e[2], o2);
        a[6] = csub<T>(e[2], o2);
        a[3] = cadd<T>(e[3], o3);
        a[7] = csub<T>(e[3], o3);
    }
};

template <typename T, int DIR> struct Dft<T, 3, DIR> {
    __device__ static inline void run(C<T>* a) {
        const T s3 = (T)0.86602540378443864676;
        C<T> t = cadd<T>(a[1], a[2]);
        C<T> d = csub<T>(a[1], a[2]);
        C<T> m = mk<T>(a[0].x - (T)0.5 * t.x, a[0].y - (T)0.5 * t.y);
        C<T> r = cmuli<T>(cscale<T>(d, s3), DIR);
        a[0] = cadd<T>(a[0], t);
        a[1] = cadd<T>(m, r);
        a[2] = csub<T>(m, r);
    }
};

// Radix 6 as two 3-point DFTs of the even and odd points, y_u = E_u + w^u O_u
// and y_{u+3} = E_u - w^u O_u (w = exp(DIR 2 pi i / 6)): one Stockham stage
// (one LDS round trip, five twiddles) where a radix-3 and a radix-2 stage
// took two (seven twiddles).
template <typename T, int DIR> struct Dft<T, 6, DIR> {
    __device__ static inline void run(C<T>* a) {
        const T s3 = (T)0.86602540378443864676;
        C<T> e[3] = {a[0], a[2], a[4]};
        C<T> o[3] = {a[1], a[3], a[5]};
        Dft<T, 3, DIR>::run(e);
        Dft<T, 3, DIR>::run(o);
        // w^1 = 1/2 + DIR i s3, w^2 = -1/2 + DIR i s3
        const C<T> o1 = mk<T>((T)0.5 * o[1].x - (T)DIR * s3 * o[1].y, (T)0.5 * o[1].y + (T)DIR * s3 * o[1].x);
        const C<T> o2 = mk<T>((T)-0.5 * o[2].x - (T)DIR * s3 * o[2].y, (T)-0.5 * o[2].y + (T)DIR * s3 * o[2].x);
        a[0] = cadd<T>(e[0], o[0]);
        a[3] = csub<T>(e[0], o[0]);
        a[1] = cadd<T>(e[1], o1);
        a[4] = csub<T>(e[1], o1);
        a[2] = cadd<T>(e[2], o2);
        a[5] = csub<T>(e[2], o2);
    }
};

template <typename T, int DIR> struct Dft<T, 5, DIR> {
    __device__ static inline void run(C<T>* a) {
        const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
        const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;
        C<T> t1 = cadd<T>(a[1], a[4]), t2 = cadd<T>(a[2], a[3]);
        C<T> d1 = csub<T>(a[1], a[4]), d2 = csub<T>(a[2], a[3]);
        C<T> m1 = mk<T>(a[0].x + c1 * t1.x + c2 * t2.x, a[0].y + c1 * t1.y + c2 * t2.y);
        C<T> m2 = mk<T>(a[0].x + c2 * t1.x + c1 * t2.x, a[0].y + c2 * t1.y + c1 * t2.y);
        C<T> r1 = cmuli<T>(mk<T>(s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y), DIR);
        C<T> r2 = cmuli<T>(mk<T>(s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y), DIR);
        a[0] = cadd<T>(a[0], cadd<T>(t1, t2));
        a[1] = cadd<T>(m1, r1);
        a[4] = csub<T>(m1, r1);
        a[2] = cadd<T>(m2, r2);
        a[3] = csub<T>(m2, r2);
    }
};

// Odd radices 7 .. 23: y_0 = sum_t a_t; for u = 1 .. (R-1)/2, with
// s_t = a_t + a_{R-t}, d_t = a_t - a_{R-t} (t = 1 .. (R-1)/2),
// y_u = a_0 + sum_t s_t cos(2 pi t u / R) + DIR i sum_t d_t sin(2 pi t u / R)
// and y_{R-u} the same with the sine term negated.  The cosines and sines are
// compile-time constants (Taylor series of the reduced angle).
constexpr double ce_sin(double x) {
    double t = x, s = x;
    for (int k = 1; k < 30; k++) {
        t *= -x * x / ((2.0 * k) * (2.0 * k + 1.0));
        s += t;
    }
    return s;
}
constexpr double ce_cos(double x) {
    double t = 1.0, s = 1.0;
    for (int k = 1; k < 30; k++) {
        t *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
        s += t;
    }
    return s;
}
template <int R> struct RootTab {
    double c[R], s[R];  // cos, sin of 2 pi m / R
    constexpr RootTab() : c(), s() {
        for (int m = 0; m < R; m++) {
            const int mm = (2 * m > R) ? m - R : m;  // angle reduced to (-pi, pi]
            const double a = 6.283185307179586476925286766559 * (double)mm / (double)R;
            c[m] = ce_cos(a);
            s[m] = ce_sin(a);
        }
    }
};
template <int R> inline constexpr RootTab<R> kRoot{};

template <typename T, int R, int DIR> struct DftOdd {
    __device__ static inline void run(C<T>* a) {
        constexpr int H = (R - 1) / 2;
        C<T> sp[H], dm[H];
#pragma unroll
        for (int t = 1; t <= H; t++) {
            sp[t - 1] = cadd<T>(a[t], a[R - t]);
            dm[t - 1] = csub<T>(a[t], a[R - t]);
        }
        C<T> y0 = a[0];
#pragma unroll
        for (int t = 0; t < H; t++) y0 = cadd<T>(y0, sp[t]);
        C<T> out[R];
        out[0] = y0;
#pragma unroll
        for (int u = 1; u <= H; u++) {
            T ax = a[0].x, ay = a[0].y, bx = (T)0, by = (T)0;
#pragma unroll
            for (int t = 1; t <= H; t++) {
                const T c = (T)kRoot<R>.c[(t * u) % R], sn = (T)kRoot<R>.s[(t * u) % R];
                ax += sp[t - 1].x * c;
                ay += sp[t - 1].y * c;
                bx += dm[t - 1].x * sn;
                by += dm[t - 1].y * sn;
            }
            // DIR i (bx + i by) = DIR (-by + i bx)
            const C<T> ib = mk<T>((T)(-DIR) * by, (T)DIR * bx);
            out[u] = mk<T>(ax + ib.x, ay + ib.y);
            out[R - u] = mk<T>(ax - ib.x, ay - ib.y);
        }
#pragma unroll
        for (int u = 0; u < R; u++) a[u] = out[u];
    }
};
template <typename T, int DIR> struct Dft<T, 7, DIR> : DftOdd<T, 7, DIR> {};
template <typename T, int DIR> struct Dft<T, 11, DIR> : DftOdd<T, 11, DIR> {};
template <typename T, int DIR> struct Dft<T, 13, DIR> : DftOdd<T, 13, DIR> {};
template <typename T, int DIR> struct Dft<T, 17, DIR> : DftOdd<T, 17, DIR> {};
template <typename T, int DIR> struct Dft<T, 19, DIR> : DftOdd<T, 19, DIR> {};
template <typename T, int DIR> struct Dft<T, 23, DIR> : DftOdd<T, 23, DIR> {};
template <typename T, int DIR> struct Dft<T, 9, DIR> : DftOdd<T, 9, DIR> {};

// ---------------------------------------------------------------------------
// Stockham stages over a tile in LDS: sm[b * LP + p], b < B, p < L
// ---------------------------------------------------------------------------
// A line of L points occupies LP slots: point p sits at slot lpos(p) =
// p + p / 32, one spare slot after every 32 points (plus one at the end).
// Eight-byte points 8 apart (the first Stockham stage writes j * 8 + t) then
// fall into 32 different bank pairs across a wavefront instead of 4.
// SW = false keeps the plain layout (slot p, LP = L + 1), which suits the
// fused triple-sum kernel's access pattern better.
template <int L, bool SW = true> struct LdsPad { static constexpr int LP = SW ? L + L / 32 + 1 : L + 1; };
template <bool SW = true> __device__ __forceinline__ int lpos(int p) { return SW ? p + (p >> 5) : p; }

template <typename T, int DIR, int L, int B, int NT, bool SW, int NS, int R, int... RS>
__device__ __forceinline__ void lds_stockham(C<T>* sm, const C<T>* __restrict__ tw, int tid) {
    constexpr int LP = LdsPad<L, SW>::LP;
    constexpr int NBF = L / R;
    constexpr int TOT = B * NBF;
    constexpr int Q = (TOT + NT - 1) / NT;
    C<T> v[Q][R];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int idx = tid + q * NT;
        if ((TOT % NT == 0) || idx < TOT) {
            const int b = idx / NBF, j = idx % NBF;
            const int k = j % NS;
#pragma unroll
            for (int t = 0; t < R; t++) {
                C<T> a = sm[b * LP + lpos<SW>(j + t * NBF)];
                if (NS > 1 && t > 0) {
                    const int e = t * k * (L / (NS * R));
                    C<T> w = tw[e];
                    if (DIR > 0) w.y = -w.y;
                    a = cmul<T>(a, w);
                }
                v[q][t] = a;
            }
            Dft<T, R, DIR>::run(v[q]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int idx = tid + q * NT;
        if ((TOT % NT == 0) || idx < TOT) {
            const int b = idx / NBF, j = idx % NBF;
            const int k = j % NS;
            const int base = (j / NS) * NS * R + k;
#pragma unroll
            for (int t = 0; t < R; t++) sm[b * LP + lpos<SW>(base + t * NS)] = v[q][t];
        }
    }
    __syncthreads();
    if constexpr (sizeof...(RS) > 0) lds_stockham<T, DIR, L, B, NT, SW, NS * R, RS...>(sm, tw, tid);
}

// Column-minor tile: point p of line b at sm[p * B + b], B a power of two
// dividing NT.  A lane keeps its line b = tid % B through every stage, the R
// points of a butterfly lie B * L / R slots apart (immediate offsets from one
// address), and the 64 lanes of a wavefront read 64 consecutive slots.  The
// index arithmetic per butterfly is the quotient and remainder of j by NS,
// where the line-major layout also divides by L / R for the line index.
template <typename T, int DIR, int L, int B, int NT, int NS, int R, int... RS>
__device__ __forceinline__ void lds_stockham_cm(C<T>* sm, const C<T>* __restrict__ tw, int tid) {
    static_assert((B & (B - 1)) == 0 && NT % B == 0, "column-minor tiles need a power-of-two width dividing NT");
    constexpr int NBF = L / R;
    constexpr int JS = NT / B;  // butterflies of one line per pass over the workgroup
    constexpr int Q = (NBF + JS - 1) / JS;
    const int b = tid % B, j0 = tid / B;
    C<T> v[Q][R];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int j = j0 + q * JS;
        if ((NBF % JS == 0) || j < NBF) {
            const int k = j % NS;
            const C<T>* src = sm + j * B + b;
#pragma unroll
            for (int t = 0; t < R; t++) {
                C<T> a = src[t * NBF * B];
                if (NS > 1 && t > 0) {
                    const int e = t * k * (L / (NS * R));
                    C<T> w = tw[e];
                    if (DIR > 0) w.y = -w.y;
                    a = cmul<T>(a, w);
                }
                v[q][t] = a;
            }
            Dft<T, R, DIR>::run(v[q]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int j = j0 + q * JS;
        if ((NBF % JS == 0) || j < NBF) {
            C<T>* dst = sm + ((j / NS) * NS * R + j % NS) * B + b;
#pragma unroll
            for (int t = 0; t < R; t++) dst[t * NS * B] = v[q][t];
        }
    }
    __syncthreads();
    if constexpr (sizeof...(RS) > 0) lds_stockham_cm<T, DIR, L, B, NT, NS * R, RS...>(sm, tw, tid);
}

// Radix plans of the supported line lengths (kFftLengths below lists them).
template <int... RS> struct Radices {};
template <int L> struct FftPlan;  // no plan: not a supported length
#define HM_FFT_PLAN(L, ...) \
    template <> struct FftPlan<L> { using type = Radices<__VA_ARGS__>; };
HM_FFT_PLAN(8, 8)
HM_FFT_PLAN(16, 4, 4)
HM_FFT_PLAN(32, 8, 4)
HM_FFT_PLAN(48, 8, 6)
HM_FFT_PLAN(64, 8, 8)
HM_FFT_PLAN(80, 8, 2, 5)
HM_FFT_PLAN(96, 8, 4, 3)
HM_FFT_PLAN(128, 8, 4, 4)
HM_FFT_PLAN(160, 8, 4, 5)
HM_FFT_PLAN(192, 8, 8, 3)
HM_FFT_PLAN(256, 8, 8, 4)
HM_FFT_PLAN(288, 8, 6, 6)
HM_FFT_PLAN(320, 8, 8, 5)
HM_FFT_PLAN(384, 8, 8, 6)
HM_FFT_PLAN(480, 8, 4, 3, 5)
HM_FFT_PLAN(512, 8, 8, 8)
HM_FFT_PLAN(576, 8, 8, 3, 3)
HM_FFT_PLAN(640, 8, 8, 2, 5)
HM_FFT_PLAN(768, 8, 8, 4, 3)
HM_FFT_PLAN(1024, 8, 8, 4, 4)
HM_FFT_PLAN(24, 8, 3)
HM_FFT_PLAN(40, 8, 5)
HM_FFT_PLAN(72, 8, 3, 3)
HM_FFT_PLAN(120, 8, 3, 5)
HM_FFT_PLAN(144, 8, 6, 3)
HM_FFT_PLAN(200, 8, 5, 5)
HM_FFT_PLAN(216, 8, 3, 3, 3)
HM_FFT_PLAN(240, 8, 6, 5)
HM_FFT_PLAN(360, 8, 3, 3, 5)
HM_FFT_PLAN(400, 8, 2, 5, 5)
HM_FFT_PLAN(432, 8, 6, 3, 3)
HM_FFT_PLAN(600, 8, 3, 5, 5)
HM_FFT_PLAN(720, 8, 6, 3, 5)
HM_FFT_PLAN(800, 8, 4, 5, 5)
HM_FFT_PLAN(960, 8, 8, 3, 5)
HM_FFT_PLAN(270, 9, 5, 6)
#undef HM_FFT_PLAN

template <typename T, int DIR, int L, int B, int NT, bool SW, int... RS>
__device__ __forceinline__ void lds_fft_run(C<T>* sm, const C<T>* __restrict__ tw, int tid, Radices<RS...>) {
    lds_stockham<T, DIR, L, B, NT, SW, 1, RS...>(sm, tw, tid);
}
template <typename T, int DIR, int L, int B, int NT, int... RS>
__device__ __forceinline__ void lds_fft_cm_run(C<T>* sm, const C<T>* __restrict__ tw, int tid, Radices<RS...>) {
    lds_stockham_cm<T, DIR, L, B, NT, 1, RS...>(sm, tw, tid);
}

// Transform of B lines of L points in LDS, line-major (sm[b * LP + lpos(p)]).
// tid: the lane's index in the workgroup (threadIdx.x unless the caller passes
// its own copy).
template <typename T, int DIR, int L, int B, int NT = FFT_NT, bool SW = true>
__device__ __forceinline__ void lds_fft(C<T>* sm, const C<T>* __restrict__ tw, int tid = -1) {
    if (tid < 0) tid = threadIdx.x;
    lds_fft_run<T, DIR, L, B, NT, SW>(sm, tw, tid, typename FftPlan<L>::type{});
}
// The same transform on a column-minor tile (sm[p * B + b]).
template <typename T, int DIR, int L, int B, int NT = FFT_NT>
__device__ __forceinline__ void lds_fft_cm(C<T>* sm, const C<T>* __restrict__ tw, int tid = -1) {
    if (tid < 0) tid = threadIdx.x;
    lds_fft_cm_run<T, DIR, L, B, NT>(sm, tw, tid, typename FftPlan<L>::type{});
}

// Lengths with a radix plan, ascending.
static constexpr int kFftLengths[] = {8, 16, 24, 32, 40, 48, 64, 72, 80, 96, 120, 128, 144, 160, 192, 200, 216, 240, 256, 270, 288, 320, 360, 384, 400, 432, 480, 512, 576, 600, 640, 720, 768, 800, 960, 1024};

constexpr int pow2floor(int x) {
    int r = 1;
    while (2 * r <= x) r *= 2;
    return r;
}

// Tile width (lines per workgroup) for a column pass: up to 128 B per point,
// a power of two, within 64 KB of LDS.  FP32 lines of 270 and 288 points take 8
// columns (64 B per point): the triple mesh's shell passes run faster with the
// smaller tiles (more workgroups in flight per CU); the autotuner can still
// choose 16 (the wide tile).
template <typename T, int L> struct ColTile {
    static constexpr int want = (sizeof(T) == 4 && (L == 288 || L == 270)) ? 8 : (int)(128 / sizeof(C<T>));
    static constexpr int fit = (int)(65536 / (LdsPad<L>::LP * sizeof(C<T>)));
    static constexpr int B = pow2floor(want < fit ? want : fit);
};
template <typename T, int L> struct RowTile {
    static constexpr int want = (4096 / L) > 0 ? (4096 / L) : 1;
    static constexpr int fit = (int)(65536 / (LdsPad<L>::LP * sizeof(C<T>)));
    static constexpr int B = pow2floor(want < fit ? want : fit);
};

// compact index -> actual index of a signed support |n| <= s on a length-n axis
__host__ __device__ inline int support_map(int c, int s, int n) {
    if (s < 0) return c;
    return (c <= s) ? c : (n - (2 * s + 1) + c);
}
__host__ __device__ inline int support_count(int s, int n) { return (s < 0 || 2 * s + 1 >= n) ? n : 2 * s + 1; }
__host__ __device__ inline bool in_support(int p, int s, int n) { return s < 0 || p <= s || p >= n - s; }

// Column-pass geometry: element (f, o, b, p) at f*FS + o*OS + b*IS + p*PS, where
// o and b are the actual outer/inner line indices and f < nf a batch index
// (grid.z) over equally laid-out arrays.
struct ColGeom {
    long long OS, PS;
    long long IS = 1;  // stride of the inner line index (1: a tile's columns are contiguous)
    long long FS = 0;
    int nf = 1;
    int n_outer, outer_full, outer_s;  // compact count, full extent, support
    int n_inner, inner_full, inner_s;
    int point_s;                       // support on the transformed axis
    // Per-batch supports: when set, batch f takes the supports of its outer,
    // inner and transformed axes from fsup[f] (x, y, z; component role_* of
    // it, -1: keep the one above).  The grid is sized for the largest; lines
    // outside a batch's own support are neither read nor written.
    const int4* fsup = nullptr;
    int role_outer = -1, role_inner = -1, role_point = -1;
    // output support on the transformed axis: only the points |n| <= out_s are stored (-1: all), for a
    // forward transform whose consumer reads a low-frequency cube only
    int out_s = -1;
    // host copy of fsup: the launcher then runs a compact grid of exactly the
    // tiles inside the batches' supports, in batch order (grid.x = all of them,
    // compact = 1), instead of the box of the largest with the rest exiting
    const int4* fsup_host = nullptr;
    int compact = 0;
};

// Row-pass geometry: element (r, p) at r*RS + p.  With fps set, row r takes
// the support fps[r / rpf] on its points instead of point_s.
struct RowGeom {
    long long n_rows, RS_in, RS_out;
    int point_s;
    const int* fps = nullptr;
    long long rpf = 1;
    int out_s = -1;  // only the output points |n| <= out_s are stored (-1: all)
};

__host__ __device__ __forceinline__ int int4_get(const int4& q, int i) { return i == 0 ? q.x : (i == 1 ? q.y : q.z); }

// Tiles of B lines of a per-batch-support column geometry (the compact grid);
// 0: the geometry has no host copy of its supports.
inline long long compact_tiles(const ColGeom& g, int B) {
    if (!g.fsup || !g.fsup_host) return 0;
    long long tot = 0;
    for (int f = 0; f < g.nf; f++) {
        const int4 q = g.fsup_host[f];
        const int no = g.role_outer >= 0 ? support_count(int4_get(q, g.role_outer), g.outer_full) : g.n_outer;
        const int ni = g.role_inner >= 0 ? support_count(int4_get(q, g.role_inner), g.inner_full) : g.n_inner;
        tot += (long long)no * ((ni + B - 1) / B);
    }
    return tot;
}

// Plain global source / destination.
template <typename T> struct GSrc {
    const C<T>* p;
    __device__ inline C<T> operator()(int /*f*/, int /*o*/, int /*b*/, int /*pt*/, long long addr) const { return p[addr]; }
};

// Workgroups are handed to the 8 XCDs round-robin by linear id, and each XCD
// has its own L2.  Renumbering the grid so that XCD x runs the x-th contiguous
// slice of linear ids keeps tiles that share 128-byte lines (neighbouring
// column tiles of 64 or 128 bytes) in one L2 instead of fetching every line
// once per XCD.  A bijection on [0, gx*gy*gz) whatever the grid.
constexpr int HM_NXCD = 8;
// (Grids here hold fewer than 2^32 workgroups.)
__device__ __forceinline__ uint3 xcd_block() {
    const unsigned gx = gridDim.x, gy = gridDim.y;
    const unsigned tot = gx * gy * gridDim.z;
    const unsigned lin = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const unsigned q = tot / HM_NXCD, r = tot % HM_NXCD;
    const unsigned x = lin % HM_NXCD, i = lin / HM_NXCD;
    const unsigned t = (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + i;
    uint3 b;
    b.x = t % gx;
    b.y = (t / gx) % gy;
    b.z = t / (gx * gy);
    return b;
}

// Column FFT: grid.x = tiles of B compact inner lines, grid.y = compact outer index.
// BW: columns per tile (0: ColTile's width).  The transform of each column is
// the same arithmetic for any width; only the tiling of the grid changes.
// A column source may split into pre(f, o, b, p) -> Raw (its loads only) and
// cook(raw, f, o, b, p) -> value (its arithmetic); k_fft_cols then issues the
// loads of all of a lane's points before cooking any.  Sources without a Raw
// type are called once per point.  The tile is column-minor in LDS (point p
// of column b at p * B + b), so a wavefront's accesses are consecutive slots.
template <class S, class = void> struct TwoPhaseSrc : std::false_type {};
template <class S> struct TwoPhaseSrc<S, std::void_t<typename S::Raw>> : std::true_type {};

template <typename T, int L, int DIR, class Src, int BW = 0>
__global__ __launch_bounds__(FFT_NT) void k_fft_cols(Src src, C<T>* __restrict__ dst, ColGeom g,
                                                    const C<T>* __restrict__ tw) {
    constexpr int B = BW ? BW : ColTile<T, L>::B;
    constexpr int LP = LdsPad<L>::LP;
    extern __shared__ unsigned char smraw[];
    C<T>* sm = reinterpret_cast<C<T>*>(smraw);
    uint3 wg = xcd_block();
    if (g.compact) {
        // compact grid: linear tile id -> (batch, compact outer index, tile) in batch order
        unsigned t = wg.x;
        int fz = -1;
        for (int ff = 0; ff < g.nf; ff++) {
            const int4 q = g.fsup[ff];
            const int no = g.role_outer >= 0 ? support_count(int4_get(q, g.role_outer), g.outer_full) : g.n_outer;
            const int ni = g.role_inner >= 0 ? support_count(int4_get(q, g.role_inner), g.inner_full) : g.n_inner;
            const unsigned nt = (unsigned)((ni + B - 1) / B), cnt = (unsigned)no * nt;
            if (t < cnt) {
                fz = ff;
                wg = make_uint3(t % nt, t / nt, (unsigned)ff);
                break;
            }
            t -= cnt;
        }
        if (fz < 0) return;
    }
    const int f = (int)wg.z;
    int outer_s = g.outer_s, inner_s = g.inner_s, point_s = g.point_s, n_inner = g.n_inner;
    if (g.fsup) {
        const int4 q = g.fsup[f];
        if (g.role_outer >= 0) {
            outer_s = int4_get(q, g.role_outer);
            if ((int)wg.y >= support_count(outer_s, g.outer_full)) return;
        }
        if (g.role_inner >= 0) {
            inner_s = int4_get(q, g.role_inner);
            n_inner = support_count(inner_s, g.inner_full);
            if ((int)wg.x * B >= n_inner) return;
        }
        if (g.role_point >= 0) point_s = int4_get(q, g.role_point);
    }
    const int o = support_map((int)wg.y, outer_s, g.outer_full);
    const long long obase = (long long)f * g.FS + (long long)o * g.OS;
    const int tid = threadIdx.x;
    const int bl = tid % B;
    const int bc = (int)wg.x * B + bl;  // compact inner index
    const bool bvalid = bc < n_inner;
    const int b = bvalid ? support_map(bc, inner_s, g.inner_full) : 0;
    constexpr int PSTEP = FFT_NT / B;
    // every point of the lane is loaded before any is written to LDS, so the
    // loads are in flight together (one at a time, each would wait its full latency)
    constexpr int NI = (L + PSTEP - 1) / PSTEP;
    const int p0 = tid / B;
    // element addresses of the lane's points: a0 + i * ast (the uniform step folds into scalar arithmetic)
    const long long a0 = obase + (long long)b * g.IS + (long long)p0 * g.PS, ast = (long long)PSTEP * g.PS;
    C<T> buf[NI];
    C<T>* stw = sm + B * LP;  // the twiddle table, staged while the tile's loads are in flight
    if constexpr (TwoPhaseSrc<Src>::value) {
        // the source's loads (pre) for every point first, its arithmetic (cook) after
        typename Src::Raw raw[NI];
        bool use[NI];
#pragma unroll
        for (int i = 0; i < NI; i++) {
            const int p = p0 + i * PSTEP;
            use[i] = p < L && bvalid && in_support(p, point_s, L);
            if (use[i]) raw[i] = src.pre(f, o, b, p);
        }
#pragma unroll
        for (int i = 0; i < NI; i++) buf[i] = use[i] ? src.cook(raw[i], f, o, b, p0 + i * PSTEP) : mk<T>((T)0, (T)0);
    } else {
#pragma unroll
        for (int i = 0; i < NI; i++) {
            const int p = p0 + i * PSTEP;
            C<T> v = mk<T>((T)0, (T)0);
            if (p < L && bvalid && in_support(p, point_s, L)) v = src(f, o, b, p, a0 + i * ast);
            buf[i] = v;
        }
    }
    for (int e = tid; e < L; e += FFT_NT) stw[e] = tw[e];
#pragma unroll
    for (int i = 0; i < NI; i++) {
        const int p = p0 + i * PSTEP;
        if (p < L) sm[p * B + bl] = buf[i];
    }
    __syncthreads();
    lds_fft_cm<T, DIR, L, B>(sm, stw);
    if (bvalid) {
        // all of the lane's points read from LDS first, then stored
#pragma unroll
        for (int i = 0; i < NI; i++) {
            const int p = p0 + i * PSTEP;
            if (p < L) buf[i] = sm[p * B + bl];
        }
#pragma unroll
        for (int i = 0; i < NI; i++) {
            const int p = p0 + i * PSTEP;
            if (p < L && in_support(p, g.out_s, L)) dst[a0 + i * ast] = buf[i];
        }
    }
}

// Row FFT over contiguous lines (complex in, complex out).
template <typename T, int L, int DIR>
__global__ __launch_bounds__(FFT_NT) void k_fft_rows(const C<T>* __restrict__ src, C<T>* __restrict__ dst, RowGeom g,
                                                    const C<T>* __restrict__ tw) {
    constexpr int B = RowTile<T, L>::B;
    constexpr int LP = LdsPad<L>::LP;
    extern __shared__ unsigned char smraw[];
    C<T>* sm = reinterpret_cast<C<T>*>(smraw);
    const long long r0 = (long long)blockIdx.x * B;
    // all of the lane's points are loaded before any is written to LDS (loads in flight together)
    constexpr int NE = (B * L + FFT_NT - 1) / FFT_NT;
    C<T> buf[NE];
#pragma unroll
    for (int i = 0; i < NE; i++) {
        const int e = threadIdx.x + i * FFT_NT;
        const int b = e / L, p = e % L;
        C<T> v = mk<T>((T)0, (T)0);
        if (e < B * L && r0 + b < g.n_rows && in_support(p, g.fps ? g.fps[(r0 + b) / g.rpf] : g.point_s, L))
            v = src[(r0 + b) * g.RS_in + p];
        buf[i] = v;
    }
    C<T>* stw = sm + B * LP;  // the twiddle table, staged while the tile's loads are in flight
    for (int e = threadIdx.x; e < L; e += FFT_NT) stw[e] = tw[e];
#pragma unroll
    for (int i = 0; i < NE; i++) {
        const int e = threadIdx.x + i * FFT_NT;
        if (e < B * L) sm[(e / L) * LP + lpos(e % L)] = buf[i];
    }
    __syncthreads();
    lds_fft<T, DIR, L, B>(sm, stw);
#pragma unroll
    for (int i = 0; i < NE; i++) {
        const int e = threadIdx.x + i * FFT_NT;
        if (e < B * L) buf[i] = sm[(e / L) * LP + lpos(e % L)];
    }
#pragma unroll
    for (int i = 0; i < NE; i++) {
        const int e = threadIdx.x + i * FFT_NT;
        const int b = e / L, p = e % L;
        if (e < B * L && r0 + b < g.n_rows && in_support(p, g.out_s, L)) dst[(r0 + b) * g.RS_out + p] = buf[i];
    }
}

// Row FFT whose input comes from a functor, src(row, point) -> complex
// (fixed-point grids converted on load); output as k_fft_rows.
template <typename T, int L, int DIR, class RSrc>
__global__ __launch_bounds__(FFT_NT) void k_fft_rows_f(RSrc src, C<T>* __restrict__ dst, RowGeom g,
                                                      const C<T>* __restrict__ tw) {
    constexpr int B = RowTile<T, L>::B;
    constexpr int LP = LdsPad<L>::LP;
    extern __shared__ unsigned char smraw[];
    C<T>* sm = reinterpret_cast<C<T>*>(smraw);
    const long long r0 = (long long)blockIdx.x * B;
    constexpr int NE = (B * L + FFT_NT - 1) / FFT_NT;
    C<T> buf[NE];
#pragma unroll
    for (int i = 0; i < NE; i++) {
        // rows past the end read the last row again (no branch around the load) and are zeroed
        const int e = threadIdx.x + i * FFT_NT;
        const int b = e / L, p = e % L;
        const bool in = ((B * L) % FFT_NT == 0 || e < B * L) && r0 + b < g.n_rows;
        const C<T> v = src(in ? r0 + b : g.n_rows - 1, in ? p : 0);
        buf[i] = in ? v : mk<T>((T)0, (T)0);
    }
    C<T>* stw = sm + B * LP;  // the twiddle table, staged while the tile's loads are in flight
    for (int e = threadIdx.x; e < L; e += FFT_NT) stw[e] = tw[e];
#pragma unroll
    for (int i = 0; i < NE; i++) {
        const int e = threadIdx.x + i * FFT_NT;
        if (e < B * L) sm[(e / L) * LP + lpos(e % L)] = buf[i];
    }
    __syncthreads();
    lds_fft<T, DIR, L, B>(sm, stw);
#pragma unroll
    for (int i = 0; i < NE; i++) {
        const int e = threadIdx.x + i * FFT_NT;
        if (e < B * L) buf[i] = sm[(e / L) * LP + lpos(e % L)];
    }
#pragma unroll
    for (int i = 0; i < NE; i++) {
        const int e = threadIdx.x + i * FFT_NT;
        const int b = e / L, p = e % L;
        if (e < B * L && r0 + b < g.n_rows && in_support(p, g.out_s, L)) dst[(r0 + b) * g.RS_out + p] = buf[i];
    }
}

// Real-to-half-complex rows: N = 2L reals per row -> L+1 complex.
// RSrc(row, m) returns the complex pair (x[2m], x[2m+1]) of a row.
// tw2 holds exp(-2 pi i k / (2L)), k < 2L.
template <typename T, int L, class RSrc>
__global__ __launch_bounds__(FFT_NT) void k_r2c_rows(RSrc src, C<T>* __restrict__ dst, long long n_rows,
                                                    const C<T>* __restrict__ tw, const C<T>* __restrict__ tw2) {
    constexpr int B = RowTile<T, L>::B;
    constexpr int LP = LdsPad<L>::LP;
    extern __shared__ unsigned char smraw[];
    C<T>* sm = reinterpret_cast<C<T>*>(smraw);
    const long long r0 = (long long)blockIdx.x * B;
    constexpr int NE = (B * L + FFT_NT - 1) / FFT_NT;
    C<T> buf[NE];
#pragma unroll
    for (int i = 0; i < NE; i++) {
        // rows past the end read the last row again (no branch around the load) and are zeroed
        const int e = threadIdx.x + i * FFT_NT;
        const int b = e / L, p = e % L;
        const bool in = ((B * L) % FFT_NT == 0 || e < B * L) && r0 + b < n_rows;
        const C<T> v = src(in ? r0 + b : n_rows - 1, in ? p : 0);
        buf[i] = in ? v : mk<T>((T)0, (T)0);
    }
    C<T>* stw = sm + B * LP;  // the twiddle table, staged while the tile's loads are in flight
    for (int e = threadIdx.x; e < L; e += FFT_NT) stw[e] = tw[e];
#pragma unroll
    for (int i = 0; i < NE; i++) {
        const int e = threadIdx.x + i * FFT_NT;
        if (e < B * L) sm[(e / L) * LP + lpos(e % L)] = buf[i];
    }
    __syncthreads();
    lds_fft<T, -1, L, B>(sm, stw);
    // X[k] = E[k] + W^k O[k], E = (Z[k] + conj Z[L-k]) / 2, O = (Z[k] - conj Z[L-k]) / (2i)
    for (int e = threadIdx.x; e < B * (L + 1); e += FFT_NT) {
        const int b = e / (L + 1), k = e % (L + 1);
        if (r0 + b >= n_rows) continue;
        const C<T> zk = sm[b * LP + lpos(k % L)];
        const C<T> zc = cconj<T>(sm[b * LP + lpos((L - k) % L)]);
        const C<T> E = mk<T>((T)0.5 * (zk.x + zc.x), (T)0.5 * (zk.y + zc.y));
        const C<T> d = mk<T>((T)0.5 * (zk.x - zc.x), (T)0.5 * (zk.y - zc.y));
        const C<T> O = mk<T>(d.y, -d.x);  // d / i
        const C<T> w = tw2[k];
        dst[(r0 + b) * (L + 1) + k] = cadd<T>(E, cmul<T>(w, O));
    }
}

// Half-complex-to-real rows (unnormalised backward): L+1 complex -> 2L reals.
// Z'[k] = (X[k] + conj X[L-k]) + i exp(+2 pi i k / 2L) (X[k] - conj X[L-k]).
template <typename T, int L, class HSrc>
__global__ __launch_bounds__(FFT_NT) void k_c2r_rows(HSrc src, T* __restrict__ dst, long long n_rows, int point_s,
                                                    const C<T>* __restrict__ tw, const C<T>* __restrict__ tw2) {
    constexpr int B = RowTile<T, L>::B;
    constexpr int LP = LdsPad<L>::LP;
    extern __shared__ unsigned char smraw[];
    C<T>* sm = reinterpret_cast<C<T>*>(smraw);
    const long long r0 = (long long)blockIdx.x * B;
    C<T>* stw = sm + B * LP;  // the twiddle table (read by every stage)
    for (int e = threadIdx.x; e < L; e += FFT_NT) stw[e] = tw[e];
    // chunks of CH elements per lane: the chunk's loads are issued before any is used
    constexpr int NE = (B * L + FFT_NT - 1) / FFT_NT, CH = NE < 8 ? NE : 8;
#pragma unroll
    for (int i0 = 0; i0 < NE; i0 += CH) {
        C<T> xks[CH], xcs[CH], tws[CH];
#pragma unroll
        for (int u = 0; u < CH; u++) {
            const int e = threadIdx.x + (i0 + u) * FFT_NT;
            const int b = e / L, k = e % L, kc = L - k;
            const bool row = (i0 + u < NE) & (e < B * L) & (r0 + b < n_rows);
            // unconditional loads at a clamped row (every row holds L + 1 points), so that
            // the chunk's loads are all in flight together; the masks (bitwise, so that no
            // branch invites the compiler to sink a load under it) select afterwards
            const long long rr = row ? r0 + b : n_rows - 1;
            const C<T> a = src(rr, k), c = src(rr, kc);
            const bool ka = row & ((point_s < 0) | (k <= point_s)), kb = row & ((point_s < 0) | (kc <= point_s));
            xks[u] = mk<T>(ka ? a.x : (T)0, ka ? a.y : (T)0);
            xcs[u] = mk<T>(kb ? c.x : (T)0, kb ? -c.y : (T)0);
            tws[u] = tw2[k];
        }
#pragma unroll
        for (int u = 0; u < CH; u++) {
            const int e = threadIdx.x + (i0 + u) * FFT_NT;
            if (i0 + u >= NE || e >= B * L) continue;
            const int b = e / L, k = e % L;
            C<T> v = mk<T>((T)0, (T)0);
            if (r0 + b < n_rows) {
                const C<T> xk = xks[u], xc = xcs[u];
                const C<T> sum = cadd<T>(xk, xc);
                const C<T> d = csub<T>(xk, xc);
                C<T> w = tws[u];
                w.y = -w.y;  // exp(+2 pi i k / 2L)
                const C<T> wd = cmul<T>(w, d);
                v = mk<T>(sum.x - wd.y, sum.y + wd.x);  // sum + i * wd
            }
            sm[b * LP + lpos(k)] = v;
        }
    }
    __syncthreads();
    lds_fft<T, +1, L, B>(sm, stw);
    for (int e = threadIdx.x; e < B * L; e += FFT_NT) {
        const int b = e / L, m = e % L;
        if (r0 + b < n_rows) {
            const C<T> z = sm[b * LP + lpos(m)];
            dst[(r0 + b) * (2 * L) + 2 * m] = z.x;
            dst[(r0 + b) * (2 * L) + 2 * m + 1] = z.y;
        }
    }
}

// Half-spectrum source for the c2r rows: row r holds L+1 complex at r*(L+1).
template <typename T> struct HalfSrc {
    const C<T>* p;
    int L;
    __device__ inline C<T> operator()(long long r, int k) const { return p[r * (L + 1) + k]; }
};

// Real rows source for the r2c rows: row r holds 2L reals.
template <typename T> struct RealSrc {
    const T* p;
    int L;
    __device__ inline C<T> operator()(long long r, int m) const {
        const C<T>* q = reinterpret_cast<const C<T>*>(p + r * (2 * (long long)L));
        return q[m];
    }
};

// LDS bytes of the three tile shapes.
// the tile, then the length's twiddle table (staged once, read by every stage)
template <typename T, int L> constexpr size_t col_lds() { return ((size_t)ColTile<T, L>::B * LdsPad<L>::LP + L) * sizeof(C<T>); }
// the wide column tile (twice ColTile's width) of the FP32 lengths the box
// plans use most; the autotuner chooses per length whether to use it
template <typename T, int L> constexpr bool has_wide_cols() {
    return sizeof(T) == 4 && (L == 256 || L == 270 || L == 288 || L == 512);
}
template <typename T, int L> constexpr size_t row_lds() { return ((size_t)RowTile<T, L>::B * LdsPad<L>::LP + L) * sizeof(C<T>); }

// ---------------------------------------------------------------------------
// Any length: the radix plan is a kernel argument.  Lengths without a
// compile-time plan (every L in [2, RT_MAX_L]) run through these kernels:
// radix 2, 3, 4, 5, 7, 8, 9, 11, 13, 17, 19 and 23 butterflies with the
// twiddle table of the length, and for any other prime factor p a direct
// p-point DFT per output (O(p) per point and stage).  A tile holds at most
// RtMax<T>::PTS points, so each lane keeps at most RtMax<T>::PTS / 256
// butterflies of a stage in registers between its LDS read and write.
// ---------------------------------------------------------------------------
constexpr int RT_MAX_L = 2048;
constexpr int RT_MAX_STAGES = 24;
template <typename T> struct RtMax { static constexpr int PTS = sizeof(T) == 4 ? 4096 : 2048; };

struct RtPlan {
    int L, LP, B, ns;
    int r[RT_MAX_STAGES];
};

// Radices of L: 8s, then 4, 2, then 3 (9s first), 5, 7, 11, 13, 17, 19, 23,
// then any remaining prime factors.  B: lines per tile (a power of two with
// B * L <= max_pts, at most want_b).  false: L outside [2, RT_MAX_L].
inline bool rt_plan_make(int L, int max_pts, int want_b, RtPlan& p) {
    if (L < 2 || L > RT_MAX_L) return false;
    p.L = L;
    p.LP = L + L / 32 + 1;
    p.ns = 0;
    int m = L;
    auto take = [&](int r) {
        while (m % r == 0 && p.ns < RT_MAX_STAGES) {
            p.r[p.ns++] = r;
            m /= r;
        }
    };
    take(8);
    take(4);
    take(2);
    take(9);
    for (int r : {3, 5, 7, 11, 13, 17, 19, 23}) take(r);
    for (int f = 29; m > 1 && f <= m; f += 2) take(f);
    if (m != 1) return false;
    int b = 1;
    while (2 * b <= want_b && 2 * b * L <= max_pts) b *= 2;
    p.B = b;
    return true;
}

__device__ __forceinline__ int rt_lpos(int p) { return p + (p >> 5); }

template <typename T, int DIR, int R, int QM>
__device__ __forceinline__ void rt_stage(C<T>* sm, const C<T>* __restrict__ tw, int L, int LP, int B, int NS, int tid) {
    const int NBF = L / R;
    const int TOT = B * NBF;
    const int tstep = L / (NS * R);
    C<T> v[QM][R];
#pragma unroll
    for (int q = 0; q < QM; q++) {
        const int idx = tid + q * FFT_NT;
        if (idx < TOT) {
            const int b = idx / NBF, j = idx - (idx / NBF) * NBF;
            const int k = j % NS;
#pragma unroll
            for (int t = 0; t < R; t++) {
                C<T> a = sm[b * LP + rt_lpos(j + t * NBF)];
                if (t > 0 && NS > 1) {
                    C<T> w = tw[t * k * tstep];
                    if (DIR > 0) w.y = -w.y;
                    a = cmul<T>(a, w);
                }
                v[q][t] = a;
            }
            Dft<T, R, DIR>::run(v[q]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < QM; q++) {
        const int idx = tid + q * FFT_NT;
        if (idx < TOT) {
            const int b = idx / NBF, j = idx - (idx / NBF) * NBF;
            const int k = j % NS;
            const int base = (j / NS) * NS * R + k;
#pragma unroll
            for (int t = 0; t < R; t++) sm[b * LP + rt_lpos(base + t * NS)] = v[q][t];
        }
    }
    __syncthreads();
}

// A stage of any radix R (a prime factor above 23): each lane forms whole
// outputs, y = sum_t (a_t w^(t k)) exp(DIR 2 pi i t u / R), from the table.
template <typename T, int DIR, int QO>
__device__ __forceinline__ void rt_stage_any(C<T>* sm, const C<T>* __restrict__ tw, int L, int LP, int B, int NS, int R,
                                             int tid) {
    const int NBF = L / R;
    const int tstep = L / (NS * R), rstep = L / R;
    const int TOT = B * L;
    C<T> out[QO];
#pragma unroll
    for (int q = 0; q < QO; q++) {
        const int o = tid + q * FFT_NT;
        C<T> acc = mk<T>((T)0, (T)0);
        if (o < TOT) {
            const int b = o / L, pp = o - (o / L) * L;
            const int k = pp % NS, u = (pp / NS) % R, j = (pp / (NS * R)) * NS + k;
            for (int t = 0; t < R; t++) {
                C<T> a = sm[b * LP + rt_lpos(j + t * NBF)];
                if (t > 0 && NS > 1) {
                    C<T> w = tw[t * k * tstep];
                    if (DIR > 0) w.y = -w.y;
                    a = cmul<T>(a, w);
                }
                C<T> w = tw[((t * u) % R) * rstep];
                if (DIR > 0) w.y = -w.y;
                acc = cadd<T>(acc, cmul<T>(a, w));
            }
        }
        out[q] = acc;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < QO; q++) {
        const int o = tid + q * FFT_NT;
        if (o < TOT) sm[(o / L) * LP + rt_lpos(o - (o / L) * L)] = out[q];
    }
    __syncthreads();
}

template <typename T, int DIR>
__device__ void rt_fft(C<T>* sm, const C<T>* __restrict__ tw, const RtPlan& pl, int tid) {
    constexpr int P = RtMax<T>::PTS;
#define HM_RT_Q(R) ((P / (R) + FFT_NT - 1) / FFT_NT)
    int NS = 1;
    for (int s = 0; s < pl.ns; s++) {
        const int R = pl.r[s];
        switch (R) {
        case 2: rt_stage<T, DIR, 2, HM_RT_Q(2)>(sm, tw, pl.L, pl.LP, pl.B, NS, tid); break;
        case 3: rt_stage<T, DIR, 3, HM_RT_Q(3)>(sm, tw, pl.L, pl.LP, pl.B, NS, tid); break;
        case 4: rt_stage<T, DIR, 4, HM_RT_Q(4)>(sm, tw, pl.L, pl.LP, pl.B, NS, tid); break;
        case 5: rt_stage<T, DIR, 5, HM_RT_Q(5)>(sm, tw, pl.L, pl.LP, pl.B, NS, tid); break;
        case 7: rt_stage<T, DIR, 7, HM_RT_Q(7)>(sm, tw, pl.L, pl.LP, pl.B, NS, tid); break;
        case 8: rt_stage<T, DIR, 8, HM_RT_Q(8)>(sm, tw, pl.L, pl.LP, pl.B, NS, tid); break;
        case 9: rt_stage<T, DIR, 9, HM_RT_Q(9)>(sm, tw, pl.L, pl.LP, pl.B, NS, tid); break;
        case 11: rt_stage<T, DIR, 11, HM_RT_Q(11)>(sm, tw, pl.L, pl.LP, pl.B, NS, tid); break;
        case 13: rt_stage<T, DIR, 13, HM_RT_Q(13)>(sm, tw, pl.L, pl.LP, pl.B, NS, tid); break;
        case 17: rt_stage<T, DIR, 17, HM_RT_Q(17)>(sm, tw, pl.L, pl.LP, pl.B, NS, tid); break;
        case 19: rt_stage<T, DIR, 19, HM_RT_Q(19)>(sm, tw, pl.L, pl.LP, pl.B, NS, tid); break;
        case 23: rt_stage<T, DIR, 23, HM_RT_Q(23)>(sm, tw, pl.L, pl.LP, pl.B, NS, tid); break;
        default: rt_stage_any<T, DIR, P / FFT_NT>(sm, tw, pl.L, pl.LP, pl.B, NS, R, tid); break;
        }
        NS *= R;
    }
#undef HM_RT_Q
}

template <typename T, int DIR, class Src>
__global__ __launch_bounds__(FFT_NT) void k_fft_cols_rt(Src src, C<T>* __restrict__ dst, ColGeom g,
                                                       const C<T>* __restrict__ tw, RtPlan pl) {
    const int B = pl.B, LP = pl.LP, L = pl.L;
    extern __shared__ unsigned char smraw[];
    C<T>* sm = reinterpret_cast<C<T>*>(smraw);
    const uint3 wg = xcd_block();
    const int f = (int)wg.z;
    int outer_s = g.outer_s, inner_s = g.inner_s, point_s = g.point_s, n_inner = g.n_inner;
    if (g.fsup) {
        const int4 q = g.fsup[f];
        if (g.role_outer >= 0) {
            outer_s = int4_get(q, g.role_outer);
            if ((int)wg.y >= support_count(outer_s, g.outer_full)) return;
        }
        if (g.role_inner >= 0) {
            inner_s = int4_get(q, g.role_inner);
            n_inner = support_count(inner_s, g.inner_full);
            if ((int)wg.x * B >= n_inner) return;
        }
        if (g.role_point >= 0) point_s = int4_get(q, g.role_point);
    }
    const int o = support_map((int)wg.y, outer_s, g.outer_full);
    const long long obase = (long long)f * g.FS + (long long)o * g.OS;
    const int tid = threadIdx.x;
    const int bl = tid % B;
    const int bc = (int)wg.x * B + bl;
    const bool bvalid = bc < n_inner;
    const int b = bvalid ? support_map(bc, inner_s, g.inner_full) : 0;
    const int pstep = FFT_NT / B;
    for (int p = tid / B; p < L; p += pstep) {
        C<T> v = mk<T>((T)0, (T)0);
        if (bvalid && in_support(p, point_s, L)) v = src(f, o, b, p, obase + (long long)b * g.IS + (long long)p * g.PS);
        sm[bl * LP + rt_lpos(p)] = v;
    }
    __syncthreads();
    rt_fft<T, DIR>(sm, tw, pl, tid);
    if (bvalid) {
        for (int p = tid / B; p < L; p += pstep)
            if (in_support(p, g.out_s, L)) dst[obase + (long long)b * g.IS + (long long)p * g.PS] = sm[bl * LP + rt_lpos(p)];
    }
}

template <typename T, int DIR>
__global__ __launch_bounds__(FFT_NT) void k_fft_rows_rt(const C<T>* __restrict__ src, C<T>* __restrict__ dst, RowGeom g,
                                                       const C<T>* __restrict__ tw, RtPlan pl) {
    const int B = pl.B, LP = pl.LP, L = pl.L;
    extern __shared__ unsigned char smraw[];
    C<T>* sm = reinterpret_cast<C<T>*>(smraw);
    const long long r0 = (long long)blockIdx.x * B;
    for (int e = threadIdx.x; e < B * L; e += FFT_NT) {
        const int b = e / L, p = e % L;
        C<T> v = mk<T>((T)0, (T)0);
        if (r0 + b < g.n_rows && in_support(p, g.fps ? g.fps[(r0 + b) / g.rpf] : g.point_s, L))
            v = src[(r0 + b) * g.RS_in + p];
        sm[b * LP + rt_lpos(p)] = v;
    }
    __syncthreads();
    rt_fft<T, DIR>(sm, tw, pl, threadIdx.x);
    for (int e = threadIdx.x; e < B * L; e += FFT_NT) {
        const int b = e / L, p = e % L;
        if (r0 + b < g.n_rows && in_support(p, g.out_s, L)) dst[(r0 + b) * g.RS_out + p] = sm[b * LP + rt_lpos(p)];
    }
}

template <typename T, int DIR, class RSrc>
__global__ __launch_bounds__(FFT_NT) void k_fft_rows_f_rt(RSrc src, C<T>* __restrict__ dst, RowGeom g,
                                                         const C<T>* __restrict__ tw, RtPlan pl) {
    const int B = pl.B, LP = pl.LP, L = pl.L;
    extern __shared__ unsigned char smraw[];
    C<T>* sm = reinterpret_cast<C<T>*>(smraw);
    const long long r0 = (long long)blockIdx.x * B;
    for (int e = threadIdx.x; e < B * L; e += FFT_NT) {
        const int b = e / L, p = e % L;
        C<T> v = mk<T>((T)0, (T)0);
        if (r0 + b < g.n_rows) v = src(r0 + b, p);
        sm[b * LP + rt_lpos(p)] = v;
    }
    __syncthreads();
    rt_fft<T, DIR>(sm, tw, pl, threadIdx.x);
    for (int e = threadIdx.x; e < B * L; e += FFT_NT) {
        const int b = e / L, p = e % L;
        if (r0 + b < g.n_rows && in_support(p, g.out_s, L)) dst[(r0 + b) * g.RS_out + p] = sm[b * LP + rt_lpos(p)];
    }
}

// k_r2c_rows and k_c2r_rows with a run-time half length L = pl.L.
template <typename T, class RSrc>
__global__ __launch_bounds__(FFT_NT) void k_r2c_rows_rt(RSrc src, C<T>* __restrict__ dst, long long n_rows,
                                                       const C<T>* __restrict__ tw, const C<T>* __restrict__ tw2,
                                                       RtPlan pl) {
    const int B = pl.B, LP = pl.LP, L = pl.L;
    extern __shared__ unsigned char smraw[];
    C<T>* sm = reinterpret_cast<C<T>*>(smraw);
    const long long r0 = (long long)blockIdx.x * B;
    for (int e = threadIdx.x; e < B * L; e += FFT_NT) {
        const int b = e / L, p = e % L;
        C<T> v = mk<T>((T)0, (T)0);
        if (r0 + b < n_rows) v = src(r0 + b, p);
        sm[b * LP + rt_lpos(p)] = v;
    }
    __syncthreads();
    rt_fft<T, -1>(sm, tw, pl, threadIdx.x);
    for (int e = threadIdx.x; e < B * (L + 1); e += FFT_NT) {
        const int b = e / (L + 1), k = e % (L + 1);
        if (r0 + b >= n_rows) continue;
        const C<T> zk = sm[b * LP + rt_lpos(k % L)];
        const C<T> zc = cconj<T>(sm[b * LP + rt_lpos((L - k) % L)]);
        const C<T> E = mk<T>((T)0.5 * (zk.x + zc.x), (T)0.5 * (zk.y + zc.y));
        const C<T> d = mk<T>((T)0.5 * (zk.x - zc.x), (T)0.5 * (zk.y - zc.y));
        const C<T> O = mk<T>(d.y, -d.x);
        const C<T> w = tw2[k];
        dst[(r0 + b) * (L + 1) + k] = cadd<T>(E, cmul<T>(w, O));
    }
}

template <typename T, class HSrc>
__global__ __launch_bounds__(FFT_NT) void k_c2r_rows_rt(HSrc src, T* __restrict__ dst, long long n_rows, int point_s,
                                                       const C<T>* __restrict__ tw, const C<T>* __restrict__ tw2,
                                                       RtPlan pl) {
    const int B = pl.B, LP = pl.LP, L = pl.L;
    extern __shared__ unsigned char smraw[];
    C<T>* sm = reinterpret_cast<C<T>*>(smraw);
    const long long r0 = (long long)blockIdx.x * B;
    for (int e = threadIdx.x; e < B * L; e += FFT_NT) {
        const int b = e / L, k = e % L;
        C<T> v = mk<T>((T)0, (T)0);
        if (r0 + b < n_rows) {
            const C<T> xk = (point_s < 0 || k <= point_s) ? src(r0 + b, k) : mk<T>((T)0, (T)0);
            const int kc = L - k;
            const C<T> xc = (point_s < 0 || kc <= point_s) ? cconj<T>(src(r0 + b, kc)) : mk<T>((T)0, (T)0);
            const C<T> sv = cadd<T>(xk, xc);
            const C<T> d = csub<T>(xk, xc);
            C<T> w = tw2[k];
            w.y = -w.y;
            const C<T> wd = cmul<T>(w, d);
            v = mk<T>(sv.x - wd.y, sv.y + wd.x);
        }
        sm[b * LP + rt_lpos(k)] = v;
    }
    __syncthreads();
    rt_fft<T, +1>(sm, tw, pl, threadIdx.x);
    for (int e = threadIdx.x; e < B * L; e += FFT_NT) {
        const int b = e / L, m = e % L;
        if (r0 + b < n_rows) {
            const C<T> z = sm[b * LP + rt_lpos(m)];
            dst[(r0 + b) * (2 * L) + 2 * m] = z.x;
            dst[(r0 + b) * (2 * L) + 2 * m + 1] = z.y;
        }
    }
}

}  // namespace hm
