#pragma once
// Device implementation of the batched Stockham FFT kernels for gfx950 (MI355X).
//
// One workgroup owns `tpb` transforms.  Each thread keeps up to EMAX complex
// points in VGPRs; a pass of radix R performs EMAX/R butterflies per thread,
// reading its inputs at stride n/R and writing them to the autosort position
// (j - k)*R + k + r*L (k = j mod L).  The first pass reads HBM directly (lane
// -contiguous float2 streams), the last pass writes HBM directly; the passes in
// between exchange through LDS (interleaved float2 points, one pad point per 16; the one-plane
// kernels below exchange real and imaginary parts in turn).
// Inverse transforms use conj(DFT(conj(x))): only forward twiddles exist.
//
// Capability it replaces: the reference's per-row Stockham device kernels
// (ops/fft/c2c/fft_regbase/fft_regbase/kernel/arch35/fft_c2c_regbase.asc:20-64,
// ops/fft/common/mixed/kernel/mixed.asc:15-31, real boundary pre/post twiddles
// ops/fft/real/fft_regbase/fft_regbase/kernel/arch35/vendor/regbase_real_boundary_impl.h:25-80),
// written here from the DFT definition for wave64 / LDS.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fft_kernel_args.h"
#include "kernels.h"
#include "radix_consts.h"

namespace asdfft {

// ------------------------------------------------------------------ helpers
template <int V>
struct IC {
    static constexpr int value = V;
};

template <int F, int... More>
struct First {
    static constexpr int value = F;
};

template <int Start, int End, typename F>
__device__ __forceinline__ void static_for(F&& f)
{
    if constexpr (Start < End) {
        f(IC<Start>{});
        static_for<Start + 1, End>(f);
    }
}

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 mul_mi(float2 a) { return make_float2(a.y, -a.x); }  // a * (-i)
__device__ __forceinline__ float2 mul_pi(float2 a) { return make_float2(-a.y, a.x); }  // a * (+i)
__device__ __forceinline__ float2 conjf2(float2 a) { return make_float2(a.x, -a.y); }

// a * e^{-2 pi i M / R} with the angle known at compile time
template <int R, int M>
__device__ __forceinline__ float2 twc(float2 x)
{
    constexpr int m = M % R;
    if constexpr (m == 0) {
        return x;
    } else if constexpr (4 * m == R) {
        return mul_mi(x);
    } else if constexpr (2 * m == R) {
        return make_float2(-x.x, -x.y);
    } else if constexpr (4 * m == 3 * R) {
        return mul_pi(x);
    } else if constexpr (8 * m == R) {
        constexpr float h = (float)RootTable<R>::c[m];
        return make_float2(h * (x.x + x.y), h * (x.y - x.x));
    } else if constexpr (8 * m == 3 * R) {
        constexpr float h = (float)RootTable<R>::s[m];
        return make_float2(h * (x.y - x.x), -h * (x.x + x.y));
    } else {
        constexpr float c = (float)RootTable<R>::c[m];
        constexpr float s = (float)RootTable<R>::s[m];
        return make_float2(x.x * c + x.y * s, x.y * c - x.x * s);
    }
}

// ------------------------------------------------------------------ codelets
// Forward DFT of R points held in v[0..R-1], natural order in and out.
template <int R>
struct Dft;

template <>
struct Dft<1> {
    static __device__ __forceinline__ void run(float2*) {}
};

template <>
struct Dft<2> {
    static __device__ __forceinline__ void run(float2* v)
    {
        float2 a = cadd(v[0], v[1]);
        v[1] = csub(v[0], v[1]);
        v[0] = a;
    }
};

template <>
struct Dft<3> {
    static __device__ __forceinline__ void run(float2* v)
    {
        constexpr float s3 = (float)RootTable<3>::s[1];  // sin(2 pi / 3)
        float2 a = cadd(v[1], v[2]);
        float2 b = csub(v[1], v[2]);
        float2 t = make_float2(v[0].x - 0.5f * a.x, v[0].y - 0.5f * a.y);
        float2 s = make_float2(b.x * s3, b.y * s3);
        v[0] = cadd(v[0], a);
        v[1] = make_float2(t.x + s.y, t.y - s.x);
        v[2] = make_float2(t.x - s.y, t.y + s.x);
    }
};

template <>
struct Dft<4> {
    static __device__ __forceinline__ void run(float2* v)
    {
        float2 t0 = cadd(v[0], v[2]);
        float2 t1 = csub(v[0], v[2]);
        float2 t2 = cadd(v[1], v[3]);
        float2 t3 = mul_mi(csub(v[1], v[3]));
        v[0] = cadd(t0, t2);
        v[2] = csub(t0, t2);
        v[1] = cadd(t1, t3);
        v[3] = csub(t1, t3);
    }
};

// Odd prime P: pairs (n, P-n) share cosines; X_k = v0 + sum a_n cos - i sum b_n sin.
template <int P>
struct DftOdd {
    static __device__ __forceinline__ void run(float2* v)
    {
        constexpr int H = (P - 1) / 2;
        float2 a[H + 1], b[H + 1];
        const float2 x0 = v[0];
        float2 sum = x0;
        static_for<1, H + 1>([&](auto I) __attribute__((always_inline)) {
            constexpr int n = decltype(I)::value;
            a[n] = cadd(v[n], v[P - n]);
            b[n] = csub(v[n], v[P - n]);
            sum = cadd(sum, a[n]);
        });
        static_for<1, H + 1>([&](auto K) __attribute__((always_inline)) {
            constexpr int k = decltype(K)::value;
            float2 re = x0;
            float2 im = make_float2(0.0f, 0.0f);
            static_for<1, H + 1>([&](auto I) __attribute__((always_inline)) {
                constexpr int n = decltype(I)::value;
                constexpr int m = (n * k) % P;
                constexpr float c = (float)RootTable<P>::c[m];
                constexpr float s = (float)RootTable<P>::s[m];
                re.x += a[n].x * c;
                re.y += a[n].y * c;
                im.x += b[n].x * s;
                im.y += b[n].y * s;
            });
            v[k] = make_float2(re.x + im.y, re.y - im.x);
            v[P - k] = make_float2(re.x - im.y, re.y + im.x);
        });
        v[0] = sum;
    }
};

template <> struct Dft<5> : DftOdd<5> {};
template <> struct Dft<7> : DftOdd<7> {};
template <> struct Dft<11> : DftOdd<11> {};
template <> struct Dft<13> : DftOdd<13> {};
template <> struct Dft<17> : DftOdd<17> {};
template <> struct Dft<19> : DftOdd<19> {};

// Composite R = A*B, decimation in time: n = A*n2 + n1, k = k1 + B*k2.
template <int A, int B>
struct DftCT {
    static __device__ __forceinline__ void run(float2* v)
    {
        constexpr int R = A * B;
        float2 y[R];
        static_for<0, A>([&](auto I1) __attribute__((always_inline)) {
            constexpr int n1 = decltype(I1)::value;
            float2 t[B];
            static_for<0, B>([&](auto I2) __attribute__((always_inline)) {
                constexpr int n2 = decltype(I2)::value;
                t[n2] = v[A * n2 + n1];
            });
            Dft<B>::run(t);
            static_for<0, B>([&](auto K1) __attribute__((always_inline)) {
                constexpr int k1 = decltype(K1)::value;
                y[n1 * B + k1] = twc<R, n1 * k1>(t[k1]);
            });
        });
        static_for<0, B>([&](auto K1) __attribute__((always_inline)) {
            constexpr int k1 = decltype(K1)::value;
            float2 t[A];
            static_for<0, A>([&](auto I1) __attribute__((always_inline)) {
                constexpr int n1 = decltype(I1)::value;
                t[n1] = y[n1 * B + k1];
            });
            Dft<A>::run(t);
            static_for<0, A>([&](auto K2) __attribute__((always_inline)) {
                constexpr int k2 = decltype(K2)::value;
                v[k1 + B * k2] = t[k2];
            });
        });
    }
};

template <> struct Dft<6> : DftCT<2, 3> {};
template <> struct Dft<8> : DftCT<2, 4> {};
template <> struct Dft<9> : DftCT<3, 3> {};
template <> struct Dft<10> : DftCT<2, 5> {};
template <> struct Dft<12> : DftCT<4, 3> {};
template <> struct Dft<16> : DftCT<4, 4> {};

// ------------------------------------------------------------------ launch geometry
// Sizes a pass needs.  The runtime kernel fills them from kernel arguments; a
// specialised kernel builds them from template constants, so after inlining every
// address below folds into immediate offsets (ds_read/ds_write offset fields,
// global_load offsets) instead of per-element VALU arithmetic.
struct Geo {
    int n;        // complex transform length
    int T;        // threads per transform
    int tpb;      // transforms per workgroup
    int tstride;  // rows mapping: padded LDS words per transform
};

__device__ __forceinline__ int pad16(int x) { return x + (x >> 4); }

// LDS slot of element i of transform tt.  Rows: one padded block per transform
// (one pad word per 16).  Columns: transforms interleaved, row pitch tpb + 1 words, so
// the slot is linear in i and lanes that take adjacent transforms hit adjacent banks.
template <int MAP>
__device__ __forceinline__ int lds_at(const Geo& g, int tt, int i)
{
    if constexpr (MAP == kMapRows) {
        return tt * g.tstride + pad16(i);
    } else {
        return i * (g.tpb + 1) + tt;
    }
}

// Slot of element base + off inside the thread's own transform block (block start
// excluded).  pbase = slot of `base`.  When off is a multiple of 16 (rows) or always
// (columns) the slot is pbase plus a constant, which a specialised kernel folds into
// the ds_read / ds_write immediate offset.  base16 asserts base % 16 == 0 and off < 16.
template <int MAP, bool RT>
__device__ __forceinline__ int lds_rel(const Geo& g, int base, int pbase, int off, bool base16)
{
    if constexpr (MAP == kMapRows) {
        if constexpr (!RT) {
            if (off % 16 == 0) return pbase + off + off / 16;
            if (base16 && off < 16) return pbase + off;
        }
        return pad16(base + off);
    } else {
        return pbase + off * (g.tpb + 1);
    }
}

template <int MAP>
__device__ __forceinline__ int lds_pbase(const Geo& g, int base)
{
    if constexpr (MAP == kMapRows) {
        return pad16(base);
    } else {
        return base * (g.tpb + 1);
    }
}

__device__ __forceinline__ float2 ldg2(const float* p, int64_t idx)
{
    return reinterpret_cast<const float2*>(p)[idx];
}

// x.y ^= sign bit when mask is 0x80000000 (conjugation without a branch)
__device__ __forceinline__ float2 conj_if(float2 x, uint32_t mask)
{
    x.y = __uint_as_float(__float_as_uint(x.y) ^ mask);
    return x;
}

// ------------------------------------------------------------------ prologue / epilogue
// Per-thread view of one transform: base pointers plus 64-bit element strides (a column
// pass of a 3D plan has stride Y*Z, and (n - 1) * stride can pass 2^31 elements).
struct Io {
    const void* in;   // first element of this transform's input
    void* out;        // first element of this transform's output
    const float* in_im;  // planar input: first element of the imaginary plane
    float* out_im;       // planar output: first element of the imaginary plane
    int64_t col;      // column index (four-step twiddle)
    uint32_t cin;     // sign mask applied to inputs (conj_in)
    uint32_t cout;    // sign mask applied to outputs (conj_out)
    int64_t in_stride;    // element stride (1 for rows)
    int64_t out_stride;
};

template <int MAP>
__device__ __forceinline__ int64_t stride_of(int64_t s)
{
    if constexpr (MAP == kMapRows) return 1;
    return s;
}

// Non-temporal load of point i of an interleaved complex input.  A multi-pass plan's first
// pass reads its input exactly once; loading it non-temporally keeps it from displacing, in the
// 256 MB device cache, the intermediate that the pass writes and the next pass reads back.
template <int MAP>
__device__ __forceinline__ float2 load_streamed(const Io& io, int i)
{
    typedef float f2v __attribute__((ext_vector_type(2)));
    const f2v q = __builtin_nontemporal_load(
        reinterpret_cast<const f2v*>(reinterpret_cast<const float2*>(io.in) + i * stride_of<MAP>(io.in_stride)));
    return make_float2(q.x, q.y);
}

template <int MAP, int PRO>
__device__ __forceinline__ float2 load_input(const FftArgs& a, const Io& io, int n, int i)
{
    const float2* in2 = reinterpret_cast<const float2*>(io.in);
    const int64_t s = stride_of<MAP>(io.in_stride);
    if constexpr (PRO == kProComplex) {
        return in2[i * s];  // conjugation (conj_in) is applied by the caller
    } else if constexpr (PRO == kProReal) {
        return make_float2(reinterpret_cast<const float*>(io.in)[i * s], 0.0f);
    } else if constexpr (PRO == kProPlanar) {
        return make_float2(reinterpret_cast<const float*>(io.in)[i * s], io.in_im[i * s]);  // conj_in by the caller
    } else if constexpr (PRO == kProShift) {
        int k = i - a.roll_in;  // element i of the shifted input is element i - roll of the input
        if (k < 0) k += n;
        return in2[k * s];      // conjugation (conj_in) is applied by the caller
    } else if constexpr (PRO == kProHerm) {
        const int h = n >> 1;
        float2 x;
        if (i <= h) {
            x = in2[i * s];
            if (i == 0 || 2 * i == n) x.y = 0.0f;
        } else {
            x = in2[(n - i) * s];
            x.y = -x.y;
        }
        return conj_if(x, io.cin);
    } else {  // kProC2REven: complex length M = n, input holds M + 1 bins
        const int m = n;
        float2 xa = conj_if(in2[i * s], io.cin);
        float2 xb = conj_if(in2[(m - i) * s], io.cin);
        if (i == 0) {
            xa.y = 0.0f;
            xb.y = 0.0f;
        }
        const float2 sm = make_float2(xa.x + xb.x, xa.y - xb.y);  // xa + conj(xb)
        const float2 d = make_float2(xa.x - xb.x, xa.y + xb.y);   // xa - conj(xb)
        const float2 w = conjf2(ldg2(a.tw_aux, i));               // W_N^{-i}
        const float2 q = cmul(w, d);
        // Z = sm + i*q ; the forward FFT consumes conj(Z)
        return make_float2(sm.x - q.y, -(sm.y + q.x));
    }
}

template <int MAP, int EPI>
__device__ __forceinline__ void store_output(const FftArgs& a, const Io& io, int n, int i, float2 x)
{
    float2* out2 = reinterpret_cast<float2*>(io.out);
    const int64_t s = stride_of<MAP>(io.out_stride);
    if constexpr (EPI == kEpiComplex) {
        out2[i * s] = x;  // conjugation (conj_out) is applied by the caller
    } else if constexpr (EPI == kEpiHalf) {
        if (2 * i <= n) out2[i * s] = x;
    } else if constexpr (EPI == kEpiReal) {
        reinterpret_cast<float*>(io.out)[i * s] = x.x;
    } else if constexpr (EPI == kEpiPlanar) {
        reinterpret_cast<float*>(io.out)[i * s] = x.x;  // conj_out applied by the caller
        io.out_im[i * s] = x.y;
    } else if constexpr (EPI == kEpiShift) {
        int k = i + a.roll_out;
        if (k >= n) k -= n;
        out2[k * s] = make_float2(x.x * a.out_scale, x.y * a.out_scale);  // conj_out applied by the caller
    } else if constexpr (EPI == kEpiTwiddle) {
        const uint32_t m = (uint32_t)io.col * (uint32_t)i;  // col * i < big length <= 2^31
        const uint32_t mask = (1u << a.big_shift) - 1u;
        const float2 w = cmul(ldg2(a.tw_big_hi, m >> a.big_shift), ldg2(a.tw_big_lo, m & mask));
        out2[i * s] = conj_if(cmul(x, w), io.cout);
    }
}

// Workgroup barrier for LDS traffic only: __syncthreads() also waits for every
// outstanding global access (vmcnt 0); here only LDS accesses (lgkmcnt) are drained.  The
// empty asm statements keep the compiler from moving LDS accesses across the barrier.
__device__ __forceinline__ void lds_barrier()
{
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

// Ordering point of a wavefront's own LDS traffic (see WS below): wait for its outstanding LDS
// accesses and keep the compiler from moving LDS accesses across.
__device__ __forceinline__ void wave_lds_sync()
{
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    __builtin_amdgcn_wave_barrier();
    asm volatile("" ::: "memory");
}

// Point `at` of the exchange buffer (interleaved float2 over both planes' space); the epilogues
// that go through LDS read the passes' results with it.
__device__ __forceinline__ float2 lds_point(const float* lre, int at)
{
    return reinterpret_cast<const float2*>(lre)[at];
}

// ------------------------------------------------------------------ one pass
// RT: runtime geometry (the pass loop dispatches on the radix at run time).  The
// opaque copies stop the compiler from hoisting every radix case's address
// arithmetic out of that loop and keeping all of it live at once (hundreds of VGPRs).
// LB: exchange barriers wait for LDS traffic only (lds_barrier), so a persistent
// workgroup's stores of the previous tile keep draining while the next tile runs.
// PRE: this pass's twiddles were loaded into registers at kernel start (twr[b * (R - 1) + r - 1]
// for butterfly b of the thread), so no table load waits inside the pass.
// WS: every transform lies inside one wavefront (rows, T divides 64) and nothing crosses
// transforms through LDS: the exchange needs no workgroup barrier.  A wavefront's LDS accesses
// complete in issue order, so its reads of a pass's inputs are done before its writes of the
// outputs land, and its writes before the next pass's reads; only the compiler has to be kept
// from reordering them (and the waves of a workgroup no longer wait for one another).
// The exchange keeps each point as one float2 (real, imaginary interleaved) in a single LDS array
// over both planes' space (lre; lim is its upper half): one 64-bit access per point instead of
// two 32-bit ones (at 4096 points 48 LDS instructions per wave instead of 128).
template <int EMAX, int R, int MAP, int PRO, int EPI, bool RT, bool LB = false, bool PRE = false, bool WS = false>
__device__ __forceinline__ void stockham_pass(const FftArgs& a, Geo g, int span, uint32_t magic, int tw_offset,
                                              bool first, bool last, float* lre, float* lim, int tt, int tid,
                                              bool live, const Io& io, const float2* twr = nullptr)
{
    float2 v[EMAX];  // pass-local: nothing is carried between passes except through LDS
    constexpr int B = EMAX / R;
    static_assert(B >= 1, "radix larger than the per-thread register budget");
    constexpr bool kEpiViaLds = (EPI == kEpiR2CEven || EPI == kEpiTranspose || EPI == kEpiTransposePlanar ||
                                 EPI == kEpiRowsOut || EPI == kEpiPairRadix2 || EPI == kEpiQuadRadix4);
    if constexpr (RT) {
        asm volatile("" : "+s"(g.n), "+s"(g.T));
        asm volatile("" : "+v"(tid));
    }
    const int n = g.n;
    const int nr = n / R;
    const int T = g.T;
    // this transform's LDS block (rows) or column (columns)
    const int toff = (MAP == kMapRows) ? tt * g.tstride : tt;
    float2* lc_t = reinterpret_cast<float2*>(lre) + toff;  // one float2 per point
    auto kmod = [&](int j) __attribute__((always_inline)) {
        if constexpr (RT) {
            return j - (int)__umulhi((uint32_t)j, magic) * span;
        } else {
            return j % span;
        }
    };

    // 1. gather
    if (first) {
        auto gather = [&](auto streamed) __attribute__((always_inline)) {
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const int j = tid + b * T;
                if (live && j < nr) {
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        if constexpr (decltype(streamed)::value) {
                            v[b * R + r] = load_streamed<MAP>(io, j + r * nr);
                        } else {
                            v[b * R + r] = load_input<MAP, PRO>(a, io, n, j + r * nr);
                        }
                    }
                }
            }
        };
        // a streamed path only in kernels that run the first launch of multi-launch plans: row
        // passes and the four-step's first (column) pass; not the twiddle-preloading kernels
        // (1D launches only) nor the other column kernels, whose code stays as it was
        if constexpr (PRO == kProComplex && !PRE && (MAP == kMapRows || EPI == kEpiTwiddle)) {
            if (a.stream_in) {  // uniform
                gather(IC<1>{});
            } else {
                gather(IC<0>{});
            }
        } else {
            gather(IC<0>{});
        }
        if constexpr (PRO == kProComplex || PRO == kProPlanar || PRO == kProShift) {
            if (io.cin) {  // uniform: one scalar branch instead of a XOR per point
#pragma unroll
                for (int e = 0; e < B * R; ++e) v[e].y = -v[e].y;
            }
        }
    } else {
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const int j = tid + b * T;
            if (j < nr) {
                const int pj = lds_pbase<MAP>(g, j);
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int at = lds_rel<MAP, RT>(g, j, pj, r * nr, false);
                    v[b * R + r] = lc_t[at];
                }
            }
        }
    }

    // 2. twiddle + butterfly
    const float2* twp = reinterpret_cast<const float2*>(a.tw) + tw_offset;
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int j = tid + b * T;
        if (j < nr) {
            if (span > 1) {
                if constexpr (PRE) {
#pragma unroll
                    for (int r = 1; r < R; ++r) v[b * R + r] = cmul(v[b * R + r], twr[b * (R - 1) + r - 1]);
                } else {
                    const int k = kmod(j);
                    const float2* w = twp + k * (R - 1);
#pragma unroll
                    for (int r = 1; r < R; ++r) v[b * R + r] = cmul(v[b * R + r], w[r - 1]);
                }
            }
            Dft<R>::run(&v[b * R]);
        }
    }

    // 3. scatter
    static_assert(!WS || (MAP == kMapRows && !kEpiViaLds), "wave-local exchange: rows, no cross-transform epilogue");
    if (!last || kEpiViaLds) {
        if constexpr (WS) {
            wave_lds_sync();
        } else if constexpr (LB) {
            lds_barrier();
        } else {
            __syncthreads();
        }
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const int j = tid + b * T;
            if (j < nr) {
                const int k = (span > 1) ? kmod(j) : 0;
                const int base = (j - k) * R + k;
                const int pb = lds_pbase<MAP>(g, base);
                const bool base16 = (span == 1) && (R % 16 == 0);  // base = j * R
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int at = lds_rel<MAP, RT>(g, base, pb, r * span, base16);
                    lc_t[at] = v[b * R + r];
                }
            }
        }
        if constexpr (WS) {
            wave_lds_sync();
        } else if constexpr (LB) {
            lds_barrier();
        } else {
            __syncthreads();
        }
    } else {
        if constexpr (EPI == kEpiComplex || EPI == kEpiHalf || EPI == kEpiPlanar || EPI == kEpiShift) {
            if (io.cout) {  // uniform branch
#pragma unroll
                for (int e = 0; e < B * R; ++e) v[e].y = -v[e].y;
            }
        }
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const int j = tid + b * T;
            if (live && j < nr) {
#pragma unroll
                for (int r = 0; r < R; ++r) store_output<MAP, EPI>(a, io, n, j + r * nr, v[b * R + r]);
            }
        }
    }
}

template <int EMAX, int R, int MAP, int PRO, int EPI, bool LB = false>
__device__ __forceinline__ void maybe_pass(const FftArgs& a, const Geo& g, int p, float* lre, float* lim, int tt,
                                           int tid, bool live, const Io& io)
{
    if constexpr (R <= EMAX) {
        const int span = __builtin_amdgcn_readfirstlane(a.pass[p].span);
        const uint32_t magic = (uint32_t)__builtin_amdgcn_readfirstlane((int)a.pass[p].magic);
        const int tw_offset = __builtin_amdgcn_readfirstlane(a.pass[p].tw_offset);
        stockham_pass<EMAX, R, MAP, PRO, EPI, true, LB>(a, g, span, magic, tw_offset, p == 0, p == a.passes - 1, lre,
                                                        lim, tt, tid, live, io);
    }
    // radices above EMAX are never scheduled: the planner checks radix <= emax before launch
}

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs on XCD b % 8),
// and each XCD has its own L2.  With remap on, XCD x takes the contiguous logical range
// [x*q + min(x, r), ...) of the g workgroups (q = g / 8, r = g % 8), so neighbouring tiles
// that share 128-byte lines (64-byte column runs) are fetched into the same L2.
__device__ __forceinline__ int64_t logical_block(int remap)
{
    const uint32_t b = blockIdx.x;
    const uint32_t g = gridDim.x;
    if (!remap || g < 16) return b;
    const uint32_t x = b & 7u, q = g >> 3, r = g & 7u;
    return (int64_t)x * q + (x < r ? x : r) + (b >> 3);
}

// q = x / d, r = x % d for x >= 0, d > 0.  Operands that fit 32 bits (every launch in practice)
// take the 32-bit division, a few instructions instead of the 64-bit expansion, which sits in
// front of every thread's first load.
__device__ __forceinline__ void divmod_nonneg(int64_t x, int64_t d, int64_t& q, int64_t& r)
{
    if ((((uint64_t)x | (uint64_t)d) >> 32) == 0) {
        const uint32_t x32 = (uint32_t)x, d32 = (uint32_t)d;
        const uint32_t q32 = x32 / d32;
        q = q32;
        r = x32 - q32 * d32;
    } else {
        q = x / d;
        r = x - q * d;
    }
}

// ------------------------------------------------------------------ shared kernel body
template <int MAP, int PRO, int EPI>
__device__ __forceinline__ void thread_setup(const FftArgs& a, const Geo& g, int& tt, int& tid, int64_t& blk, bool& live,
                                             Io& io)
{
    const int lane = threadIdx.x;
    if constexpr (MAP == kMapRows) {
        tt = lane / g.T;
        tid = lane - tt * g.T;
    } else {
        tt = lane % g.tpb;
        tid = lane / g.tpb;
    }
    blk = logical_block(a.xcd_remap);
    const int64_t t = blk * g.tpb + tt;
    live = t < a.count;
    io.in = a.in;
    io.out = a.out;
    io.in_im = a.in_im;
    io.out_im = a.out_im;
    io.col = 0;
    io.cin = a.conj_in ? 0x80000000u : 0u;
    io.cout = a.conj_out ? 0x80000000u : 0u;
    io.in_stride = a.in_stride;
    io.out_stride = a.out_stride;
    if (live) {
        int64_t o = t;  // outer index; one transform per outer step when cols == 1 (uniform)
        if (a.cols != 1) divmod_nonneg(t, a.cols, o, io.col);
        int64_t oi = o;     // outer index read (input shift: rolled within groups of in_o_mod)
        int64_t co = io.col;  // column index stored (output shift: rolled)
        if constexpr (PRO == kProShift) {
            if (a.in_o_mod > 1) {
                const int64_t d = o % a.in_o_mod;
                int64_t e = d - a.in_o_roll;
                if (e < 0) e += a.in_o_mod;
                oi = o - d + e;
            }
        }
        if constexpr (EPI == kEpiShift) {
            co += a.out_col_roll;
            if (co >= a.cols) co -= a.cols;
        }
        int64_t ib = io.col * a.in_inner + oi * a.in_outer;
        int64_t ob = co * a.out_inner + o * a.out_outer;
        if (a.o_split > 0) {  // uniform
            int64_t iq, ir, oq, orr;
            divmod_nonneg(oi, a.o_split, iq, ir);
            divmod_nonneg(o, a.o_split, oq, orr);
            ib = io.col * a.in_inner + ir * a.in_outer + iq * a.in_outer2;
            ob = co * a.out_inner + orr * a.out_outer + oq * a.out_outer2;
        }
        if constexpr (PRO == kProReal) {
            io.in = reinterpret_cast<const float*>(a.in) + ib;
        } else if constexpr (PRO == kProPlanar) {
            io.in = reinterpret_cast<const float*>(a.in) + ib;
            io.in_im = a.in_im + ib;
        } else {
            io.in = reinterpret_cast<const float2*>(a.in) + ib;
        }
        if constexpr (EPI == kEpiReal) {
            io.out = reinterpret_cast<float*>(a.out) + ob;
        } else if constexpr (EPI == kEpiPlanar) {
            io.out = reinterpret_cast<float*>(a.out) + ob;
            io.out_im = a.out_im + ob;
        } else {
            io.out = reinterpret_cast<float2*>(a.out) + ob;
        }
    }
}

template <int MAP, int EPI>
__device__ __forceinline__ void lds_epilogue(const FftArgs& a, const Geo& g, float* lre, float* lim, int tt, int tid,
                                             int64_t blk, bool live, const Io& io)
{
    if constexpr (EPI == kEpiR2CEven) {
        // LDS holds Z = FFT_M(x[2n] + i x[2n+1]) in natural order.
        // X[k] = 1/2 [(Z_k + conj Z_{M-k}) - i W_N^k (Z_k - conj Z_{M-k})], k in [0, M]
        if (live) {
            const int m = g.n;
            float2* out2 = reinterpret_cast<float2*>(io.out);
            for (int k = tid; 2 * k <= m; k += g.T) {
                const int km = (k == 0) ? 0 : m - k;
                const int at0 = lds_at<MAP>(g, tt, k);
                const int at1 = lds_at<MAP>(g, tt, km);
                const float2 zk = lds_point(lre, at0);
                const float2 zm = lds_point(lre, at1);
                const float2 wk = ldg2(a.tw_aux, k);
                const float2 wm = ldg2(a.tw_aux, m - k);
                const float2 s0 = make_float2(zk.x + zm.x, zk.y - zm.y);
                const float2 d0 = make_float2(zk.x - zm.x, zk.y + zm.y);
                const float2 q0 = cmul(wk, d0);
                const float2 x0 = make_float2(0.5f * (s0.x + q0.y), 0.5f * (s0.y - q0.x));
                const float2 s1 = make_float2(zm.x + zk.x, zm.y - zk.y);
                const float2 d1 = make_float2(zm.x - zk.x, zm.y + zk.y);
                const float2 q1 = cmul(wm, d1);
                const float2 x1 = make_float2(0.5f * (s1.x + q1.y), 0.5f * (s1.y - q1.x));
                out2[k] = conj_if(x0, io.cout);
                if (m - k != k) out2[m - k] = conj_if(x1, io.cout);
            }
        }
    } else if constexpr (EPI == kEpiTranspose || EPI == kEpiTransposePlanar) {
        // LDS holds tpb finished transforms; adjacent lanes take adjacent transforms so the
        // strided destination (transform t -> column) is written in contiguous runs.
        const int lane = threadIdx.x;
        const int tt2 = lane % g.tpb;
        const int step = g.T;
        const int64_t t2 = blk * g.tpb + tt2;
        if (t2 < a.count) {
            const int64_t c2 = t2 % a.cols;
            const int64_t o = c2 * a.out_inner + (t2 / a.cols) * a.out_outer;
            const int64_t s = a.out_stride;
            if constexpr (EPI == kEpiTranspose) {
                float2* ob = reinterpret_cast<float2*>(a.out) + o;
                for (int i = lane / g.tpb; i < g.n; i += step) {
                    const int at = lds_at<MAP>(g, tt2, i);
                    ob[i * s] = conj_if(lds_point(lre, at), io.cout);
                }
            } else {
                float* obr = reinterpret_cast<float*>(a.out) + o;
                float* obi = a.out_im + o;
                const float sgn = io.cout ? -1.0f : 1.0f;
                for (int i = lane / g.tpb; i < g.n; i += step) {
                    const int at = lds_at<MAP>(g, tt2, i);
                    const float2 x = lds_point(lre, at);
                    obr[i * s] = x.x;
                    obi[i * s] = sgn * x.y;
                }
            }
        }
    } else if constexpr (EPI == kEpiPairRadix2) {
        // rows mapping, two transforms per workgroup: 2 blk (row n) and 2 blk + 1 (row n + X/2
        // of the same slice), n = blk % o_split.  First radix-2 stage of the length-X column
        // transforms (decimation in frequency): u + v is stored at n out_outer + slice out_outer2
        // and (u - v) W_X^n out_inner further on (the planner interleaves them as rows 2n and
        // 2n + 1), W_X^n = tw_aux[n] (forward sign; the inverse comes from the conjugations)
        const int64_t t0 = blk * 2;
        if (t0 + 1 < a.count) {
            const float2 w = ldg2(a.tw_aux, blk % a.o_split);
            const int64_t ob = (blk % a.o_split) * a.out_outer + (blk / a.o_split) * a.out_outer2;
            float2* out0 = reinterpret_cast<float2*>(a.out) + ob;
            float2* out1 = out0 + a.out_inner;
            for (int i = threadIdx.x; i < g.n; i += g.T * 2) {
                const int at0 = lds_at<MAP>(g, 0, i), at1 = lds_at<MAP>(g, 1, i);
                const float2 u = lds_point(lre, at0);
                const float2 v = lds_point(lre, at1);
                out0[i] = conj_if(cadd(u, v), io.cout);
                out1[i] = conj_if(cmul(csub(u, v), w), io.cout);
            }
        }
    } else if constexpr (EPI == kEpiQuadRadix4) {
        // rows mapping, four transforms per workgroup: 4 blk + q is row n + q X/4 of one slice,
        // n = blk % o_split.  First radix-4 stage of the length-X column transforms (decimation
        // in frequency): y_k = sum_q u_q W_4^(q k), stored times W_X^(n k) (tw_aux[n k]) at
        // n out_outer + slice out_outer2 + k out_inner (the planner makes that row 4n + k)
        const int64_t t0 = blk * 4;
        if (t0 + 3 < a.count) {
            const int64_t nn = blk % a.o_split;
            const float2 w1 = ldg2(a.tw_aux, nn), w2 = ldg2(a.tw_aux, 2 * nn), w3 = ldg2(a.tw_aux, 3 * nn);
            const int64_t ob = nn * a.out_outer + (blk / a.o_split) * a.out_outer2;
            float2* o0 = reinterpret_cast<float2*>(a.out) + ob;
            float2* o1 = o0 + a.out_inner;
            float2* o2 = o1 + a.out_inner;
            float2* o3 = o2 + a.out_inner;
            for (int i = threadIdx.x; i < g.n; i += g.T * 4) {
                const int at0 = lds_at<MAP>(g, 0, i), at1 = lds_at<MAP>(g, 1, i);
                const int at2 = lds_at<MAP>(g, 2, i), at3 = lds_at<MAP>(g, 3, i);
                const float2 u0 = lds_point(lre, at0), u1 = lds_point(lre, at1);
                const float2 u2 = lds_point(lre, at2), u3 = lds_point(lre, at3);
                const float2 a0 = cadd(u0, u2), a1 = csub(u0, u2);
                const float2 b0 = cadd(u1, u3), b1 = mul_mi(csub(u1, u3));
                o0[i] = conj_if(cadd(a0, b0), io.cout);
                o1[i] = conj_if(cmul(cadd(a1, b1), w1), io.cout);
                o2[i] = conj_if(cmul(csub(a0, b0), w2), io.cout);
                o3[i] = conj_if(cmul(csub(a1, b1), w3), io.cout);
            }
        }
    } else if constexpr (EPI == kEpiRowsOut) {
        // column mapping: LDS holds the workgroup's tpb transforms interleaved; each is written
        // as one contiguous row (out_stride 1), consecutive lanes on consecutive elements
        const int nthreads = g.T * g.tpb;
        const int total = g.tpb * g.n;
        for (int e = threadIdx.x; e < total; e += nthreads) {
            const int tt2 = e / g.n;
            const int i = e - tt2 * g.n;
            const int64_t t2 = blk * g.tpb + tt2;
            if (t2 >= a.count) break;  // e only grows: the rest of this thread's elements are past the end
            const int64_t o = (t2 % a.cols) * a.out_inner + (t2 / a.cols) * a.out_outer;
            const int at = lds_at<MAP>(g, tt2, i);
            reinterpret_cast<float2*>(a.out)[o + i] = conj_if(lds_point(lre, at), io.cout);
        }
    }
}

template <int EMAX>
constexpr int max_block()
{
    return EMAX <= 16 ? 256 : 512;
}

// Occupancy target of kernels that run EMAX-16 transforms in 512-thread workgroups: at
// most 128 VGPRs (4 waves per SIMD); EMAX-32 kernels hold twice the points.
template <int EMAX>
constexpr int min_waves()
{
    return EMAX <= 16 ? 4 : 1;
}


// ------------------------------------------------------------------ runtime kernel
template <int EMAX, int MAP, int PRO, int EPI>
__global__ void __launch_bounds__(max_block<EMAX>()) stockham_kernel(const FftArgs a)
{
    extern __shared__ float smem[];
    float* lre = smem;
    float* lim = smem + a.lds_plane;
    const Geo g{a.n, a.threads, a.tpb, a.lds_tstride};
    int tt, tid;
    int64_t blk;
    bool live;
    Io io;
    thread_setup<MAP, PRO, EPI>(a, g, tt, tid, blk, live, io);

    for (int p = 0; p < a.passes; ++p) {
        // pass descriptors are uniform; readfirstlane keeps the switch a scalar branch
        switch (__builtin_amdgcn_readfirstlane(a.pass[p].radix)) {
#define ASDFFT_CASE(RR) \
    case RR: maybe_pass<EMAX, RR, MAP, PRO, EPI>(a, g, p, lre, lim, tt, tid, live, io); break;
            ASDFFT_CASE(16)
            ASDFFT_CASE(12)
            ASDFFT_CASE(10)
            ASDFFT_CASE(9)
            ASDFFT_CASE(8)
            ASDFFT_CASE(6)
            ASDFFT_CASE(4)
            ASDFFT_CASE(2)
            ASDFFT_CASE(19)
            ASDFFT_CASE(17)
            ASDFFT_CASE(13)
            ASDFFT_CASE(11)
            ASDFFT_CASE(7)
            ASDFFT_CASE(5)
            ASDFFT_CASE(3)
#undef ASDFFT_CASE
            default:
                break;
        }
    }
    lds_epilogue<MAP, EPI>(a, g, lre, lim, tt, tid, blk, live, io);
}

// ------------------------------------------------------------------ specialised kernel
// Same passes with N, threads, transforms per workgroup and the radix sequence as
// template constants (instances listed in static_instances.h).
// Row transforms that fit inside one wavefront with an epilogue that stays inside the transform.
template <int MAP, int EPI, int T>
constexpr bool wave_local_exchange()
{
    return MAP == kMapRows && T <= 64 && 64 % T == 0 &&
           (EPI == kEpiComplex || EPI == kEpiHalf || EPI == kEpiReal || EPI == kEpiPlanar || EPI == kEpiShift);
}


template <int EMAX, int MAP, int PRO, int EPI, int N, int T, int TPB, int P, int SPAN, int TWOFF, int R, int... Rest>
__device__ __forceinline__ void static_passes(const FftArgs& a, const Geo& g, float* lre, float* lim, int tt, int tid,
                                              bool live, const Io& io)
{
    stockham_pass<EMAX, R, MAP, PRO, EPI, false, false, false, wave_local_exchange<MAP, EPI, T>()>(
        a, g, SPAN, 0u, TWOFF, P == 0, sizeof...(Rest) == 0, lre, lim, tt, tid, live, io);
    if constexpr (sizeof...(Rest) > 0) {
        constexpr int kNextOff = TWOFF + (SPAN > 1 ? SPAN * (R - 1) : 0);
        static_passes<EMAX, MAP, PRO, EPI, N, T, TPB, P + 1, SPAN * R, kNextOff, Rest...>(a, g, lre, lim, tt, tid,
                                                                                         live, io);
    }
}

// Column kernels know their workgroup (T x TPB) at compile time and take it as the launch
// bound, so an EMAX-16 column tile may use 512 threads; row kernels keep tpb at run time, and
// a row instance wider than the register budget's block (one transform per workgroup) takes
// its own thread count, rounded up to whole wavefronts.
template <int EMAX, int MAP, int T, int TPB>
constexpr int static_block_bound()
{
    // columns, and row instances with a fixed number of transforms per workgroup
    if constexpr (MAP == kMapCols || TPB > 1) return T * TPB;
    return T > max_block<EMAX>() ? (T + 63) / 64 * 64 : max_block<EMAX>();
}

template <int EMAX, int MAP, int PRO, int EPI, int N, int T, int TPB, int... Rs>
__global__ void __launch_bounds__((static_block_bound<EMAX, MAP, T, TPB>())) stockham_static_kernel(const FftArgs a)
{
    extern __shared__ float smem[];
    constexpr int kTstride = N + (N >> 4) + 1;
    float* lre = smem;
    float* lim = smem + a.lds_plane;
    // rows mapping keeps tpb at run time (it only offsets the transform's LDS block);
    // the column mapping interleaves transforms inside LDS, so there it is a constant
    const Geo g{N, T, (MAP == kMapRows) ? a.tpb : TPB, kTstride};
    int tt, tid;
    int64_t blk;
    bool live;
    Io io;
    thread_setup<MAP, PRO, EPI>(a, g, tt, tid, blk, live, io);
    static_passes<EMAX, MAP, PRO, EPI, N, T, TPB, 0, 1, 0, Rs...>(a, g, lre, lim, tt, tid, live, io);
    lds_epilogue<MAP, EPI>(a, g, lre, lim, tt, tid, blk, live, io);
}

// ------------------------------------------------------------------ twiddles preloaded
// Specialised row kernels for small launches (a few hundred workgroups, latency-bound): every
// pass's twiddles are loaded into registers at kernel start, beside the input loads, instead
// of one dependent table load per pass after each LDS exchange.  Pass p of radix R takes
// (EMAX / R) * (R - 1) entries (none for the first pass, whose span is 1).
template <int EMAX, int... Rs>
constexpr int tw_preload_count()
{
    int span = 1, total = 0;
    for (int r : {Rs...}) {
        if (span > 1) total += (EMAX / r) * (r - 1);
        span *= r;
    }
    return total;
}

template <int EMAX, int N, int T, int SPAN, int TWOFF, int R, int... Rest>
__device__ __forceinline__ void preload_twiddles(const FftArgs& a, int tid, float2* twr)
{
    constexpr int B = EMAX / R;
    constexpr int NR = N / R;
    if constexpr (SPAN > 1) {
        const float2* twp = reinterpret_cast<const float2*>(a.tw) + TWOFF;
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const int j = tid + b * T;
            const int k = (j < NR) ? j % SPAN : 0;  // threads past the pass's range read entry 0
#pragma unroll
            for (int r = 1; r < R; ++r) twr[b * (R - 1) + r - 1] = twp[k * (R - 1) + r - 1];
        }
    }
    if constexpr (sizeof...(Rest) > 0) {
        constexpr int kNextOff = TWOFF + (SPAN > 1 ? SPAN * (R - 1) : 0);
        preload_twiddles<EMAX, N, T, SPAN * R, kNextOff, Rest...>(a, tid, twr + (SPAN > 1 ? B * (R - 1) : 0));
    }
}

template <int EMAX, int MAP, int PRO, int EPI, int N, int T, int TPB, int P, int SPAN, int TWOFF, int R, int... Rest>
__device__ __forceinline__ void static_passes_pre(const FftArgs& a, const Geo& g, float* lre, float* lim, int tt,
                                                  int tid, bool live, const Io& io, const float2* twr)
{
    stockham_pass<EMAX, R, MAP, PRO, EPI, false, false, true>(a, g, SPAN, 0u, TWOFF, P == 0, sizeof...(Rest) == 0,
                                                             lre, lim, tt, tid, live, io, twr);
    if constexpr (sizeof...(Rest) > 0) {
        constexpr int kNextOff = TWOFF + (SPAN > 1 ? SPAN * (R - 1) : 0);
        static_passes_pre<EMAX, MAP, PRO, EPI, N, T, TPB, P + 1, SPAN * R, kNextOff, Rest...>(
            a, g, lre, lim, tt, tid, live, io, twr + (SPAN > 1 ? (EMAX / R) * (R - 1) : 0));
    }
}

template <int EMAX, int MAP, int PRO, int EPI, int N, int T, int TPB, int... Rs>
__global__ void __launch_bounds__((static_block_bound<EMAX, MAP, T, TPB>())) stockham_static_pre_kernel(const FftArgs a)
{
    extern __shared__ float smem[];
    constexpr int kTstride = N + (N >> 4) + 1;
    float* lre = smem;
    float* lim = smem + a.lds_plane;
    const Geo g{N, T, (MAP == kMapRows) ? a.tpb : TPB, kTstride};
    int tt, tid;
    int64_t blk;
    bool live;
    Io io;
    thread_setup<MAP, PRO, EPI>(a, g, tt, tid, blk, live, io);
    constexpr int kPre = tw_preload_count<EMAX, Rs...>();
    float2 twr[kPre > 0 ? kPre : 1];
    preload_twiddles<EMAX, N, T, 1, 0, Rs...>(a, tid, twr);
    static_passes_pre<EMAX, MAP, PRO, EPI, N, T, TPB, 0, 1, 0, Rs...>(a, g, lre, lim, tt, tid, live, io, twr);
    lds_epilogue<MAP, EPI>(a, g, lre, lim, tt, tid, blk, live, io);
}

// ------------------------------------------------------------------ one-plane exchange
// Occupancy target of the one-plane row kernels.  A 512-thread EMAX-16 row workgroup (8
// waves; the 2D rows in fours) takes about 34 KiB of LDS, so four fit a CU by LDS; four need 8
// waves per SIMD, at most 64 VGPRs.  Left to itself the compiler takes 72 VGPRs there (7 waves
// per SIMD: three workgroups per CU).  With the target, and each butterfly's twiddles loaded
// only after the previous butterfly is done (otherwise the compiler spills to scratch to meet
// it), the 2048-point rows-in-fours pass of 2048^2 x 4 took 59.0-59.3 against 60.9-61.8 us
// (the whole transform stays at 108-109 us: its two passes' 537 MB of reads and writes in
// 108 us are already the 5.0 TB/s of a plain device copy).  The 512-thread one-plane column
// kernels measured slower both ways (512-point columns: 66.1 staged, 67.1 staged at the
// target, against 64.4 us), and 32 columns per 1024-thread workgroup at the target (56 VGPRs,
// two workgroups per CU) no faster (65.0), so the column kernels keep the compiler's choice.
template <int EMAX, int MAP, int T, int TPB>
constexpr int split_min_waves()
{
    return (EMAX <= 16 && MAP == kMapRows && static_block_bound<EMAX, MAP, T, TPB>() == 512) ? 8 : 1;
}

// Specialised kernel whose passes exchange through ONE LDS plane: first the real parts of
// every point, then the imaginary parts.  It needs half the LDS of the two-plane exchange,
// so twice the workgroups fit a CU and one's loads overlap the others' passes; the cost is
// two more barriers per exchange.  Columns: long columns (2048 points, 8 columns: 72 KiB
// instead of 144 KiB) run two workgroups per CU instead of one.  Rows: 256-thread
// workgroups of EMAX-16 transforms (about 17 KiB instead of 34) reach the 8 waves per SIMD
// that their register use allows instead of 4.  The points stay in registers across passes
// (the exchange overwrites them plane by plane: all real parts are stored before any is
// read back, and the imaginary parts, still in the previous pass's layout, are stored only
// after that).
template <int EMAX, int MAP, int EPI, int N, int T, int TPB, int P, int SPAN, int TWOFF, int R, int... Rest>
__device__ __forceinline__ void split_passes(const FftArgs& a, const Geo& g, float* lp, int tt, int tid, bool live,
                                             const Io& io, float2 (&v)[EMAX], int64_t blk)
{
    constexpr int B = EMAX / R;
    constexpr int NR = N / R;
    constexpr int PITCH = TPB + 1;  // column layout: slot of element i of column tt = i * PITCH + tt
    constexpr int kTstride = N + (N >> 4) + 1;  // rows layout: one padded block per transform
    constexpr bool kStaged = split_min_waves<EMAX, MAP, T, TPB>() > 1;
    static_assert(B >= 1, "radix larger than the per-thread register budget");
    float* lp_t = (MAP == kMapRows) ? lp + tt * kTstride : lp + tt;
    if constexpr (P == 0) {
        const float2* in2 = reinterpret_cast<const float2*>(io.in);
        const int64_t s = stride_of<MAP>(io.in_stride);
        auto gather = [&](auto streamed) __attribute__((always_inline)) {
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const int j = tid + b * T;
                if (live && j < NR) {
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        if constexpr (decltype(streamed)::value) {
                            v[b * R + r] = load_streamed<MAP>(io, j + r * NR);
                        } else {
                            v[b * R + r] = in2[(j + r * NR) * s];
                        }
                    }
                }
            }
        };
        if constexpr (MAP == kMapRows) {  // the 2D rows in fours run first launches
            if (a.stream_in) {  // uniform
                gather(IC<1>{});
            } else {
                gather(IC<0>{});
            }
        } else {
            gather(IC<0>{});
        }
        if (io.cin) {
#pragma unroll
            for (int e = 0; e < B * R; ++e) v[e].y = -v[e].y;
        }
    }
    const float2* twp = reinterpret_cast<const float2*>(a.tw) + TWOFF;
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int j = tid + b * T;
        if (j < NR) {
            if constexpr (SPAN > 1) {
                const float2* w = twp + (j % SPAN) * (R - 1);
#pragma unroll
                for (int r = 1; r < R; ++r) v[b * R + r] = cmul(v[b * R + r], w[r - 1]);
            }
            Dft<R>::run(&v[b * R]);
        }
        if constexpr (SPAN > 1 && kStaged) {
            // butterfly b is done before the next one's twiddles load: one butterfly's
            // twiddles in registers at a time keeps the kernel inside its occupancy target
#pragma unroll
            for (int r = 0; r < R; ++r) asm volatile("" : "+v"(v[b * R + r].x), "+v"(v[b * R + r].y)::"memory");
        }
    }
    if constexpr (sizeof...(Rest) == 0 && EPI == kEpiQuadRadix4) {
        // rows in fours (kEpiQuadRadix4 through one plane): the four rows' outputs go through
        // the plane once per component, read back so that each thread holds element i of all
        // four rows for i = threadIdx.x + m 4T; then the columns' first radix-4 stage in
        // registers, as in lds_epilogue
        static_assert(MAP == kMapRows, "rows in fours use the rows mapping");
        constexpr int kM = (N + 4 * T - 1) / (4 * T);
        float2 c[4 * kM];  // c[4 m + q]: element threadIdx.x + m 4T of row q
        static_for<0, 2>([&](auto PL) __attribute__((always_inline)) {
            constexpr bool kIm = decltype(PL)::value == 1;
            __syncthreads();  // the previous plane's reads are done before it is overwritten
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const int j = tid + b * T;
                if (j < NR) {
#pragma unroll
                    for (int r = 0; r < R; ++r) lp_t[pad16(j + r * NR)] = kIm ? v[b * R + r].y : v[b * R + r].x;
                }
            }
            __syncthreads();
#pragma unroll
            for (int m = 0; m < kM; ++m) {
                const int i = (int)threadIdx.x + m * 4 * T;
                if (i < N) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const float x = lp[q * kTstride + pad16(i)];
                        if constexpr (kIm) {
                            c[4 * m + q].y = x;
                        } else {
                            c[4 * m + q].x = x;
                        }
                    }
                }
            }
        });
        if (blk * 4 + 3 < a.count) {
            const int64_t nn = blk % a.o_split;
            const float2 w1 = ldg2(a.tw_aux, nn), w2 = ldg2(a.tw_aux, 2 * nn), w3 = ldg2(a.tw_aux, 3 * nn);
            const int64_t ob = nn * a.out_outer + (blk / a.o_split) * a.out_outer2;
            float2* o0 = reinterpret_cast<float2*>(a.out) + ob;
            float2* o1 = o0 + a.out_inner;
            float2* o2 = o1 + a.out_inner;
            float2* o3 = o2 + a.out_inner;
#pragma unroll
            for (int m = 0; m < kM; ++m) {
                const int i = (int)threadIdx.x + m * 4 * T;
                if (i < N) {
                    const float2 u0 = c[4 * m], u1 = c[4 * m + 1], u2 = c[4 * m + 2], u3 = c[4 * m + 3];
                    const float2 a0 = cadd(u0, u2), a1 = csub(u0, u2);
                    const float2 b0 = cadd(u1, u3), b1 = mul_mi(csub(u1, u3));
                    o0[i] = conj_if(cadd(a0, b0), io.cout);
                    o1[i] = conj_if(cmul(cadd(a1, b1), w1), io.cout);
                    o2[i] = conj_if(cmul(csub(a0, b0), w2), io.cout);
                    o3[i] = conj_if(cmul(csub(a1, b1), w3), io.cout);
                }
            }
        }
    } else if constexpr (sizeof...(Rest) == 0) {
        if (io.cout) {
#pragma unroll
            for (int e = 0; e < B * R; ++e) v[e].y = -v[e].y;
        }
        float2* out2 = reinterpret_cast<float2*>(io.out);
        const int64_t s = stride_of<MAP>(io.out_stride);
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const int j = tid + b * T;
            if (live && j < NR) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    if constexpr (EPI == kEpiShift) {
                        store_output<MAP, kEpiShift>(a, io, N, j + r * NR, v[b * R + r]);
                    } else {
                        out2[(j + r * NR) * s] = v[b * R + r];
                    }
                }
            }
        }
    } else {
        constexpr int RN = First<Rest...>::value;  // the next pass's radix
        constexpr int BN = EMAX / RN;
        constexpr int NRN = N / RN;
        static_for<0, 2>([&](auto PL) __attribute__((always_inline)) {
            constexpr bool kIm = decltype(PL)::value == 1;
            __syncthreads();  // the previous plane's reads are done before it is overwritten
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const int j = tid + b * T;
                if (j < NR) {
                    const int k = (SPAN > 1) ? j % SPAN : 0;
                    const int base = (j - k) * R + k;
                    const int pb = lds_pbase<MAP>(g, base);
                    constexpr bool kBase16 = (SPAN == 1) && (R % 16 == 0);  // base = j * R
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        const int at = (MAP == kMapRows) ? lds_rel<kMapRows, false>(g, base, pb, r * SPAN, kBase16)
                                                         : (base + r * SPAN) * PITCH;
                        lp_t[at] = kIm ? v[b * R + r].y : v[b * R + r].x;
                    }
                }
            }
            __syncthreads();
#pragma unroll
            for (int b = 0; b < BN; ++b) {
                const int j = tid + b * T;
                if (j < NRN) {
                    const int pj = lds_pbase<MAP>(g, j);
#pragma unroll
                    for (int r = 0; r < RN; ++r) {
                        const int at = (MAP == kMapRows) ? lds_rel<kMapRows, false>(g, j, pj, r * NRN, false)
                                                         : (j + r * NRN) * PITCH;
                        const float x = lp_t[at];
                        if constexpr (kIm) {
                            v[b * RN + r].y = x;
                        } else {
                            v[b * RN + r].x = x;
                        }
                    }
                }
            }
        });
        constexpr int kNextOff = TWOFF + (SPAN > 1 ? SPAN * (R - 1) : 0);
        split_passes<EMAX, MAP, EPI, N, T, TPB, P + 1, SPAN * R, kNextOff, Rest...>(a, g, lp, tt, tid, live, io, v,
                                                                                     blk);
    }
}

// Rows mapping keeps transforms per workgroup at run time (the launch bound is the
// register-budget block); columns take it as a constant.
template <int EMAX, int MAP, int EPI, int N, int T, int TPB, int... Rs>
__global__ void __launch_bounds__((static_block_bound<EMAX, MAP, T, TPB>()))
__attribute__((amdgpu_waves_per_eu(split_min_waves<EMAX, MAP, T, TPB>()))) stockham_split_kernel(const FftArgs a)
{
    extern __shared__ float smem[];
    const Geo g{N, T, (MAP == kMapRows) ? a.tpb : TPB, N + (N >> 4) + 1};
    int tt, tid;
    int64_t blk;
    bool live;
    Io io;
    thread_setup<MAP, kProComplex, EPI>(a, g, tt, tid, blk, live, io);
    float2 v[EMAX];
    split_passes<EMAX, MAP, EPI, N, T, TPB, 0, 1, 0, Rs...>(a, g, smem, tt, tid, live, io, v, blk);
}

// ------------------------------------------------------------------ launch helper
template <int V> struct VariantTraits;
template <> struct VariantTraits<kVarRowsC2C> { static constexpr int map = kMapRows, pro = kProComplex, epi = kEpiComplex; };
template <> struct VariantTraits<kVarColsC2C> { static constexpr int map = kMapCols, pro = kProComplex, epi = kEpiComplex; };
template <> struct VariantTraits<kVarR2CEven> { static constexpr int map = kMapRows, pro = kProComplex, epi = kEpiR2CEven; };
template <> struct VariantTraits<kVarC2REven> { static constexpr int map = kMapRows, pro = kProC2REven, epi = kEpiComplex; };
template <> struct VariantTraits<kVarR2COdd> { static constexpr int map = kMapRows, pro = kProReal, epi = kEpiHalf; };
template <> struct VariantTraits<kVarC2ROdd> { static constexpr int map = kMapRows, pro = kProHerm, epi = kEpiReal; };
template <> struct VariantTraits<kVarColsTwiddle> { static constexpr int map = kMapCols, pro = kProComplex, epi = kEpiTwiddle; };
template <> struct VariantTraits<kVarRowsTranspose> { static constexpr int map = kMapRows, pro = kProComplex, epi = kEpiTranspose; };
template <> struct VariantTraits<kVarRowsSep> { static constexpr int map = kMapRows, pro = kProPlanar, epi = kEpiPlanar; };
template <> struct VariantTraits<kVarRowsSepIn> { static constexpr int map = kMapRows, pro = kProPlanar, epi = kEpiComplex; };
template <> struct VariantTraits<kVarColsSepOut> { static constexpr int map = kMapCols, pro = kProComplex, epi = kEpiPlanar; };
template <> struct VariantTraits<kVarColsTwiddleSep> { static constexpr int map = kMapCols, pro = kProPlanar, epi = kEpiTwiddle; };
template <> struct VariantTraits<kVarRowsTransposeSep> { static constexpr int map = kMapRows, pro = kProComplex, epi = kEpiTransposePlanar; };
template <> struct VariantTraits<kVarColsToRows> { static constexpr int map = kMapCols, pro = kProComplex, epi = kEpiRowsOut; };
template <> struct VariantTraits<kVarRowsShiftIn> { static constexpr int map = kMapRows, pro = kProShift, epi = kEpiComplex; };
template <> struct VariantTraits<kVarRowsShiftBoth> { static constexpr int map = kMapRows, pro = kProShift, epi = kEpiShift; };
template <> struct VariantTraits<kVarColsShiftOut> { static constexpr int map = kMapCols, pro = kProComplex, epi = kEpiShift; };
template <> struct VariantTraits<kVarRowsPair2> { static constexpr int map = kMapRows, pro = kProComplex, epi = kEpiPairRadix2; };
template <> struct VariantTraits<kVarRowsC2CTwPre> { static constexpr int map = kMapRows, pro = kProComplex, epi = kEpiComplex; };
template <> struct VariantTraits<kVarRowsQuad4> { static constexpr int map = kMapRows, pro = kProComplex, epi = kEpiQuadRadix4; };

// One instantiation per kernel, so the LDS limit is raised once per kernel and device.
template <auto Kern>
hipError_t launch_kernel(const LaunchShape& s, const FftArgs& args, hipStream_t stream)
{
    if (s.lds_bytes > 65536) {
        static std::atomic<uint64_t> raised{0};
        hipError_t e = raise_lds_limit_once(reinterpret_cast<const void*>(Kern), raised);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(Kern, dim3((unsigned)s.grid), dim3((unsigned)s.block), s.lds_bytes, stream, args);
    return hipGetLastError();
}

template <int EMAX, int V>
hipError_t launch_variant(const LaunchShape& s, const FftArgs& args, hipStream_t stream)
{
    using Tr = VariantTraits<V>;
    return launch_kernel<stockham_kernel<EMAX, Tr::map, Tr::pro, Tr::epi>>(s, args, stream);
}

template <int V, int EMAX, int N, int T, int TPB, int... Rs>
hipError_t launch_static(const LaunchShape& s, const FftArgs& args, hipStream_t stream)
{
    if constexpr (V == kVarColsC2CSplit) {
        return launch_kernel<stockham_split_kernel<EMAX, kMapCols, kEpiComplex, N, T, TPB, Rs...>>(s, args, stream);
    } else if constexpr (V == kVarColsShiftOutSplit) {
        return launch_kernel<stockham_split_kernel<EMAX, kMapCols, kEpiShift, N, T, TPB, Rs...>>(s, args, stream);
    } else if constexpr (V == kVarRowsC2CSplit) {
        return launch_kernel<stockham_split_kernel<EMAX, kMapRows, kEpiComplex, N, T, TPB, Rs...>>(s, args, stream);
    } else if constexpr (V == kVarRowsQuad4Split) {
        return launch_kernel<stockham_split_kernel<EMAX, kMapRows, kEpiQuadRadix4, N, T, TPB, Rs...>>(s, args, stream);
    } else if constexpr (V == kVarRowsC2CTwPre) {
        return launch_kernel<stockham_static_pre_kernel<EMAX, kMapRows, kProComplex, kEpiComplex, N, T, TPB, Rs...>>(
            s, args, stream);
    } else {
        using Tr = VariantTraits<V>;
        return launch_kernel<stockham_static_kernel<EMAX, Tr::map, Tr::pro, Tr::epi, N, T, TPB, Rs...>>(s, args,
                                                                                                       stream);
    }
}

}  // namespace asdfft
