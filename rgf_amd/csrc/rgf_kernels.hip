// gfx950 kernels of the RGF slice recursion.
//
//  K1 assemble_slice   M = A0 - k^2 diag(eps_n) - G_{n-1}: one dense pass
//                      (HBM-bound) plus a CSR pass for the stencil A0.
//  K2 invert_inplace   blocked complex Gauss-Jordan inverse; the rank-b
//                      trailing updates run on MFMA through cgemm.
//  K3 cgemm            complex GEMM on MFMA (v_mfma_f64_16x16x4 /
//                      v_mfma_f32_16x16x4), planar re/im operands, 4M or 3M.
//  K4 fill_modes       transverse mode matrices for the lead Green's function
//                      (lead_green_separable: the same in two contractions
//                      for a 3D lead), fill_mode_cols: the selected mode
//                      columns of the lead projections.
//  K5 (projection)     Fisher-Lee products are cgemm calls from the driver.
//  residual_stencil    FP64 residual of the block-tridiagonal system used by
//                      the mixed-precision refinement.
//
// Complex matrices are planar (separate re and im planes) and row-major, so
// every MFMA operand is a plain real fragment and the four (or three) real
// products of a complex product are independent MFMA chains.

#include "rgf_internal.h"

#include <hip/hip_ext.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <type_traits>

namespace rgf {

static int grid_for(int64_t total, int threads = 256, int cap = 16384) {
    return (int)std::max<int64_t>(1, std::min<int64_t>((total + threads - 1) / threads, cap));
}

// Rows [rlo, rhi), columns [clo, chi) and the diagonal block
// [blo, bhi) x [blo, bhi) of C that a GEMM leaves untouched.
struct Skip {
    int64_t rlo = 0, rhi = 0, clo = 0, chi = 0, blo = 0, bhi = 0;
};

// ------------------------------------------------------------------ MFMA wrappers
template <typename T>
struct Mfma;

template <>
struct Mfma<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mma(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // accumulator element r of lane l holds row (l/16) + 4r
    static __device__ __forceinline__ int row(int l4, int r) { return l4 + 4 * r; }
    // lane group and element holding row k
    static constexpr __device__ __forceinline__ int grp(int k) { return k % 4; }
    static constexpr __device__ __forceinline__ int elt(int k) { return k / 4; }
};

template <>
struct Mfma<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mma(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // accumulator element r of lane l holds row 4(l/16) + r
    static __device__ __forceinline__ int row(int l4, int r) { return 4 * l4 + r; }
    // lane group and element holding row k
    static constexpr __device__ __forceinline__ int grp(int k) { return k / 4; }
    static constexpr __device__ __forceinline__ int elt(int k) { return k % 4; }
};

// ------------------------------------------------------------------ K3: complex GEMM
// Work-group tile BM x BN, K chunk BK, WM x WN wavefronts; each wavefront owns
// a (BM/WM) x (BN/WN) sub-tile made of 16x16 MFMA blocks.  MFMA operand
// layout (16x16x4): lane l supplies A[l%16][l/16] and B[l/16][l%16]; the
// accumulator element r of lane l is D[Mfma<T>::row(l/16, r)][l%16], which is
// 4(l/16)+r for f32 and (l/16)+4r for f64 (measured on gfx950).
//
// USE3M: Re = ArBr - AiBi, Im = (Ar+Ai)(Br+Bi) - ArBr - AiBi (three MFMA
// chains instead of four; the operand sums are formed once per LDS tile).
template <typename T>
struct Vec;  // 16-byte vector of T for the interior-tile global loads
template <>
struct Vec<float> {
    typedef float type __attribute__((ext_vector_type(4)));
    static constexpr int W = 4;
};
template <>
struct Vec<double> {
    typedef double type __attribute__((ext_vector_type(2)));
    static constexpr int W = 2;
};

// One GEMM problem (planar operands, row-major).
template <typename T>
struct GemmArgs {
    const T* Are;
    const T* Aim;
    int64_t lda;
    const T* Bre;
    const T* Bim;
    int64_t ldb;
    T* Cre;
    T* Cim;
    int64_t ldc;
    const T* Xre;  // beta = 1 source of C (C itself unless set otherwise)
    const T* Xim;
    int64_t ldx;
    // optional second destination: C[i][j] for i in [y_rlo, y_rhi), j in
    // [y_clo, y_chi) is also stored to Y[i - y_rlo][j - y_clo]
    T* Yre;
    T* Yim;
    int64_t ldy;
    int y_rlo, y_rhi, y_clo, y_chi;
    int M, N, K;
    double alpha_re, alpha_im;
    int beta;
    int srlo, srhi, sclo, schi;
    int sblo, sbhi;  // skipped diagonal block
    int vec_ok;
    int nz, kc;       // split-K: nz chunks of kc along K, chunk z writes C + z * zstride
    int64_t zstride;
    int tiles_n, tiles;  // tile grid of this problem (set at launch)
    // compressed tile grid (MAP bit 2): tile rows [cr_lo, cr_lo + cr_n) and
    // tile columns [cc_lo, cc_lo + cc_n), which the skips cover whole, are
    // left out of the grid (set at launch)
    int cr_lo, cr_n, cc_lo, cc_n;
    // optional: max |C_ij|^2 over the stored entries, as the bits of a
    // non-negative float (which order like unsigned integers), atomicMax'ed
    unsigned* amax;
    // optional hand-over between the problems of one launch (work groups are
    // dispatched in grid order, so a waiting problem must come after the
    // problem it waits for): a tile whose columns meet [ctr_clo, ctr_chi)
    // adds 1 to *ctr after its stores are visible device-wide; a problem with
    // ctr_wait > 0 waits until *ctr reaches it before reading its operands
    // (bounded: on a time-out bit 2 of *fail is set and the tile goes on)
    unsigned* ctr;
    int ctr_clo, ctr_chi;
    unsigned ctr_wait;
    int* fail;
};

// Spin limit of a hand-over wait (2^20 polls of the counter, each a
// device-scope load and a 64-cycle sleep: at most about a second; the tiles
// waited for take tens of microseconds).
constexpr unsigned kHandoverSpins = 1u << 20;

// max |x|^2 of a wavefront's values into *amax (non-negative float bits)
__device__ __forceinline__ void wave_amax(unsigned* amax, float m) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    // only a value above the current maximum needs the atomic (almost every
    // wavefront skips it: one address takes all the updates of a launch)
    const unsigned v = __float_as_uint(isnan(m) ? INFINITY : m);
    if ((threadIdx.x & 63) == 0 && v > __hip_atomic_load(amax, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        atomicMax(amax, v);
}

// A 1D grid covers the tiles of g0 followed by those of g1: two independent
// problems sharing the tile configuration run as one launch.
template <typename T, int BM, int BN, int BK, int WM, int WN, bool USE3M>
__global__ __launch_bounds__(WM* WN * 64) void cgemm_kernel(GemmArgs<T> g0, GemmArgs<T> g1) {
    const bool second = (int)blockIdx.x >= g0.tiles;
    const GemmArgs<T>& g = second ? g1 : g0;
    int lin = second ? (int)blockIdx.x - g0.tiles : (int)blockIdx.x;
    int z = 0;
    if (g.nz > 1) {
        const int per = g.tiles / g.nz;
        z = lin / per;
        lin -= z * per;
    }
    const int64_t koff = (int64_t)z * g.kc;
    const int64_t lda = g.lda;
    const int64_t ldb = g.ldb;
    const T* __restrict__ Are = g.Are + koff;
    const T* __restrict__ Aim = g.Aim + koff;
    const T* __restrict__ Bre = g.Bre + koff * ldb;
    const T* __restrict__ Bim = g.Bim + koff * ldb;
    T* Cre = g.Cre + z * g.zstride;
    T* Cim = g.Cim + z * g.zstride;
    const int64_t ldc = g.ldc;
    const int M = g.M, N = g.N;
    const int K = g.nz > 1 ? (int)min((int64_t)g.kc, (int64_t)g.K - koff) : g.K;
    const double alpha_re = g.alpha_re, alpha_im = g.alpha_im;
    const int beta = g.beta;
    const int srlo = g.srlo, srhi = g.srhi, sclo = g.sclo, schi = g.schi;
    const int sblo = g.sblo, sbhi = g.sbhi;
    const int vec_ok = g.vec_ok;
    using V = typename Mfma<T>::v4;
    using VT = typename Vec<T>::type;
    constexpr int VW = Vec<T>::W;
    constexpr int NT = WM * WN * 64;
    constexpr int TM = BM / WM, TN = BN / WN;
    constexpr int FM = TM / 16, FN = TN / 16;
    constexpr int NP = USE3M ? 3 : 2;  // LDS planes: re, im (, re+im)
    __shared__ T sA[NP][BM][BK + 1];
    __shared__ T sB[NP][BK][BN + 1];

    const int tile_n = (lin % g.tiles_n) * BN;
    const int tile_m = (lin / g.tiles_n) * BM;
    if (tile_m >= M || tile_n >= N) return;
    // tiles lying entirely inside the skipped band exit; partially covered
    // tiles are masked element-wise in the epilogue
    {
        const bool rs = tile_m >= srlo && tile_m + BM <= srhi;
        const bool cs = tile_n >= sclo && tile_n + BN <= schi;
        const bool bs = tile_m >= sblo && tile_m + BM <= sbhi && tile_n >= sblo && tile_n + BN <= sbhi;
        if (rs || cs || bs) return;
    }
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave / WN, wn = wave % WN;
    const int l16 = lane & 15, l4 = lane >> 4;

    // beta = 1 with alpha = +-1: the accumulators start from C (loads issued
    // before the K loop) and the sign of alpha is folded into the A operand.
    const bool accc = beta && alpha_im == 0.0 && (alpha_re == 1.0 || alpha_re == -1.0);
    const T sgn = accc ? (T)alpha_re : T(1);
    V acc0[FM][FN], acc1[FM][FN], acc2[FM][FN];
#pragma unroll
    for (int a = 0; a < FM; ++a)
#pragma unroll
        for (int b = 0; b < FN; ++b) {
            acc0[a][b] = V{0, 0, 0, 0};
            acc1[a][b] = V{0, 0, 0, 0};
            acc2[a][b] = V{0, 0, 0, 0};
            if (accc) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int gi = tile_m + wm * TM + a * 16 + Mfma<T>::row(l4, r);
                    const int gj = tile_n + wn * TN + b * 16 + l16;
                    if (gi < M && gj < N) {
                        const int64_t o = (int64_t)gi * g.ldx + gj;
                        const T cr = g.Xre[o], ci = g.Xim[o];
                        acc0[a][b][r] = cr;
                        if (USE3M)
                            acc2[a][b][r] = cr + ci;  // Im = acc2 - acc0 - acc1 keeps C_im
                        else
                            acc1[a][b][r] = ci;
                    }
                }
            }
        }

    // Interior tiles with 16-byte aligned rows load whole vectors without
    // bounds checks; edge tiles load element by element.
    const bool interior = vec_ok && tile_m + BM <= M && tile_n + BN <= N && (K % BK) == 0;
    constexpr int AE = BM * BK, BE = BK * BN;
    constexpr int AQ = (AE + NT - 1) / NT, BQ = (BE + NT - 1) / NT;
    constexpr int AV = AE / VW, BV = BE / VW;
    constexpr int AQV = (AV + NT - 1) / NT, BQV = (BV + NT - 1) / NT;
    T pa_r[AQ], pa_i[AQ], pb_r[BQ], pb_i[BQ];
    VT va_r[AQV], va_i[AQV], vb_r[BQV], vb_i[BQV];
    const T* Ar0 = Are + (int64_t)tile_m * lda;
    const T* Ai0 = Aim + (int64_t)tile_m * lda;
    auto fetch = [&](int k0) {
        if (interior) {
#pragma unroll
            for (int q = 0; q < AQV; ++q) {
                const int e = tid + q * NT;
                if (e < AV) {
                    const int i = e / (BK / VW), kv = (e % (BK / VW)) * VW;
                    const int o = i * (int)lda + k0 + kv;
                    va_r[q] = *(const VT*)(Ar0 + o);
                    va_i[q] = *(const VT*)(Ai0 + o);
                }
            }
#pragma unroll
            for (int q = 0; q < BQV; ++q) {
                const int e = tid + q * NT;
                if (e < BV) {
                    const int k = e / (BN / VW), jv = (e % (BN / VW)) * VW;
                    const int64_t o = (int64_t)(k0 + k) * ldb + tile_n + jv;
                    vb_r[q] = *(const VT*)(Bre + o);
                    vb_i[q] = *(const VT*)(Bim + o);
                }
            }
            return;
        }
#pragma unroll
        for (int q = 0; q < AQ; ++q) {
            const int e = tid + q * NT;
            const int i = e / BK, k = e % BK;
            const int gi = tile_m + i, gk = k0 + k;
            pa_r[q] = 0;
            pa_i[q] = 0;
            if (e < AE && gi < M && gk < K) {
                pa_r[q] = Are[(int64_t)gi * lda + gk];
                pa_i[q] = Aim[(int64_t)gi * lda + gk];
            }
        }
#pragma unroll
        for (int q = 0; q < BQ; ++q) {
            const int e = tid + q * NT;
            const int k = e / BN, j = e % BN;
            const int gk = k0 + k, gj = tile_n + j;
            pb_r[q] = 0;
            pb_i[q] = 0;
            if (e < BE && gk < K && gj < N) {
                pb_r[q] = Bre[(int64_t)gk * ldb + gj];
                pb_i[q] = Bim[(int64_t)gk * ldb + gj];
            }
        }
    };
    auto stash = [&]() {
        if (interior) {
#pragma unroll
            for (int q = 0; q < AQV; ++q) {
                const int e = tid + q * NT;
                if (e < AV) {
                    const int i = e / (BK / VW), kv = (e % (BK / VW)) * VW;
#pragma unroll
                    for (int u = 0; u < VW; ++u) {
                        sA[0][i][kv + u] = va_r[q][u];
                        sA[1][i][kv + u] = va_i[q][u];
                        if (USE3M) sA[NP - 1][i][kv + u] = va_r[q][u] + va_i[q][u];
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < BQV; ++q) {
                const int e = tid + q * NT;
                if (e < BV) {
                    const int k = e / (BN / VW), jv = (e % (BN / VW)) * VW;
#pragma unroll
                    for (int u = 0; u < VW; ++u) {
                        sB[0][k][jv + u] = vb_r[q][u];
                        sB[1][k][jv + u] = vb_i[q][u];
                        if (USE3M) sB[NP - 1][k][jv + u] = vb_r[q][u] + vb_i[q][u];
                    }
                }
            }
            return;
        }
#pragma unroll
        for (int q = 0; q < AQ; ++q) {
            const int e = tid + q * NT;
            if (e < AE) {
                const int i = e / BK, k = e % BK;
                sA[0][i][k] = pa_r[q];
                sA[1][i][k] = pa_i[q];
                if (USE3M) sA[NP - 1][i][k] = pa_r[q] + pa_i[q];
            }
        }
#pragma unroll
        for (int q = 0; q < BQ; ++q) {
            const int e = tid + q * NT;
            if (e < BE) {
                const int k = e / BN, j = e % BN;
                sB[0][k][j] = pb_r[q];
                sB[1][k][j] = pb_i[q];
                if (USE3M) sB[NP - 1][k][j] = pb_r[q] + pb_i[q];
            }
        }
    };

    fetch(0);
    for (int k0 = 0; k0 < K; k0 += BK) {
        stash();
        __syncthreads();
        if (k0 + BK < K) fetch(k0 + BK);
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            T ar[FM], ai[FM], nai[FM], br[FN], bi[FN];
#pragma unroll
            for (int a = 0; a < FM; ++a) {
                const int r = wm * TM + a * 16 + l16;
                ar[a] = sgn * sA[0][r][kk + l4];
                ai[a] = sgn * sA[1][r][kk + l4];
                nai[a] = -ai[a];
            }
#pragma unroll
            for (int b = 0; b < FN; ++b) {
                const int c = wn * TN + b * 16 + l16;
                br[b] = sB[0][kk + l4][c];
                bi[b] = sB[1][kk + l4][c];
            }
            // independent MFMA sweeps: a chain's dependent MFMAs are
            // FM*FN (or 2*FM*FN) instructions apart
            if (USE3M) {
                T as[FM], bs[FN];
#pragma unroll
                for (int a = 0; a < FM; ++a) as[a] = sgn * sA[NP - 1][wm * TM + a * 16 + l16][kk + l4];
#pragma unroll
                for (int b = 0; b < FN; ++b) bs[b] = sB[NP - 1][kk + l4][wn * TN + b * 16 + l16];
#pragma unroll
                for (int a = 0; a < FM; ++a)
#pragma unroll
                    for (int b = 0; b < FN; ++b) acc0[a][b] = Mfma<T>::mma(ar[a], br[b], acc0[a][b]);
#pragma unroll
                for (int a = 0; a < FM; ++a)
#pragma unroll
                    for (int b = 0; b < FN; ++b) acc1[a][b] = Mfma<T>::mma(ai[a], bi[b], acc1[a][b]);
#pragma unroll
                for (int a = 0; a < FM; ++a)
#pragma unroll
                    for (int b = 0; b < FN; ++b) acc2[a][b] = Mfma<T>::mma(as[a], bs[b], acc2[a][b]);
            } else {
#pragma unroll
                for (int a = 0; a < FM; ++a)
#pragma unroll
                    for (int b = 0; b < FN; ++b) {
                        acc0[a][b] = Mfma<T>::mma(ar[a], br[b], acc0[a][b]);
                        acc1[a][b] = Mfma<T>::mma(ar[a], bi[b], acc1[a][b]);
                    }
#pragma unroll
                for (int a = 0; a < FM; ++a)
#pragma unroll
                    for (int b = 0; b < FN; ++b) {
                        acc0[a][b] = Mfma<T>::mma(nai[a], bi[b], acc0[a][b]);
                        acc1[a][b] = Mfma<T>::mma(ai[a], br[b], acc1[a][b]);
                    }
            }
        }
        __syncthreads();
    }

    // epilogue
    float mloc = 0.f;
#pragma unroll
    for (int a = 0; a < FM; ++a)
#pragma unroll
        for (int b = 0; b < FN; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gi = tile_m + wm * TM + a * 16 + Mfma<T>::row(l4, r);
                const int gj = tile_n + wn * TN + b * 16 + l16;
                const bool masked = (gi >= srlo && gi < srhi) || (gj >= sclo && gj < schi) ||
                                    (gi >= sblo && gi < sbhi && gj >= sblo && gj < sbhi);
                if (gi < M && gj < N && !masked) {
                    T pr, pi;
                    if (USE3M) {
                        pr = acc0[a][b][r] - acc1[a][b][r];
                        pi = acc2[a][b][r] - acc0[a][b][r] - acc1[a][b][r];
                    } else {
                        pr = acc0[a][b][r];
                        pi = acc1[a][b][r];
                    }
                    const int64_t o = (int64_t)gi * ldc + gj;
                    if (!accc) {
                        double vr = alpha_re * (double)pr - alpha_im * (double)pi;
                        double vi = alpha_re * (double)pi + alpha_im * (double)pr;
                        if (beta) {
                            const int64_t ox = (int64_t)gi * g.ldx + gj;
                            vr += (double)g.Xre[ox];
                            vi += (double)g.Xim[ox];
                        }
                        pr = (T)vr;
                        pi = (T)vi;
                    }
                    Cre[o] = pr;
                    Cim[o] = pi;
                    mloc = fmaxf(mloc, (float)(pr * pr + pi * pi));
                    if (gi >= g.y_rlo && gi < g.y_rhi && gj >= g.y_clo && gj < g.y_chi) {
                        const int64_t oy = (int64_t)(gi - g.y_rlo) * g.ldy + (gj - g.y_clo);
                        g.Yre[oy] = pr;
                        g.Yim[oy] = pi;
                    }
                }
            }
    if (g.amax) wave_amax(g.amax, mloc);
}

// Streamlined tile kernel for alpha = +-1, beta in {0, 1}, K a multiple of
// BK and 16-byte aligned operand rows (the shapes of the slice inverse and
// the refinement sweeps).  Differences to cgemm_kernel:
//  * loads are unconditional: rows beyond M and column vectors beyond N read
//    clamped in-bounds addresses (B rows must hold N rounded up to the vector
//    width), the epilogue masks them;
//  * LDS is double-buffered (one barrier per K chunk);
//  * the K index of MFMA step s at lane group g is 4g + s, so each lane
//    takes its four A values of a 16-deep K block with one 16-byte LDS read;
//  * ACC3: three accumulator sets (ArBr, AiBi, ArBi + AiBr), no negated
//    operands; otherwise two sets with -Ai formed in registers (FP64, where
//    the accumulators dominate the register budget);
//  * the four MFMA sweeps of a K step are ordered so that dependent MFMAs
//    are FM*FN instructions apart;
//  * the sign of alpha is applied when A is staged into LDS;
//  * MAP 1: the linear index is remapped so that consecutive tiles stay on
//    one XCD (work groups are dealt round-robin over the 8 XCDs) and tiles
//    run column-major in groups of 8 tile rows; MAP 0 keeps the row-major
//    order with round-robin XCDs (measured equal or faster: the operand
//    panels of the shapes used here stay in L2 either way);
//  * MAP bit 2: tile rows and columns that the skips cover whole are left
//    out of the grid instead of launching work groups that return at once
//    (the slice-inverse bulk update: 3596 work groups instead of 4096 at
//    N = 4096, b = 128; 123 us instead of 139-147 us alone).
// One output tile (linear index lin of g's tile grid, split-K chunk
// included) of the streamlined kernel; the whole work group takes part.
// AMAX: also track max |C_ij|^2 into g.amax (the column panels of the
// slice inverse; a separate instantiation keeps the registers of the other
// products, the trailing update among them, at four waves per SIMD).
template <typename T, int BM, int BN, int BK, int WM, int WN, bool ACC3, int MAP, bool TR, bool M3, bool AMAX = false>
__device__ __forceinline__ void cgemm2_tile(const GemmArgs<T>& g, int lin) {
    int z = 0;
    if (g.nz > 1) {
        const int per = g.tiles / g.nz;
        z = lin / per;
        lin -= z * per;
    }
    const int64_t koff = (int64_t)z * g.kc;
    const int64_t lda = g.lda, ldb = g.ldb, ldc = g.ldc;
    const T* __restrict__ Are = g.Are + koff;
    const T* __restrict__ Aim = g.Aim + koff;
    const T* __restrict__ Bre = g.Bre + koff * ldb;
    const T* __restrict__ Bim = g.Bim + koff * ldb;
    T* Cre = g.Cre + z * g.zstride;
    T* Cim = g.Cim + z * g.zstride;
    const int M = g.M, N = g.N;
    const int K = g.nz > 1 ? (int)min((int64_t)g.kc, (int64_t)g.K - koff) : g.K;
    int tile_n, tile_m;
    if constexpr ((MAP & 1) != 0) {
        // groups of 8 tile rows, column-major inside a group
        const int tiles_m = g.tiles / g.nz / g.tiles_n;
        const int grp = lin / (8 * g.tiles_n), first = grp * 8;
        const int gr = min(8, tiles_m - first), idx = lin - grp * 8 * g.tiles_n;
        tile_m = first + idx % gr;
        tile_n = idx / gr;
    } else {
        tile_n = lin % g.tiles_n;
        tile_m = lin / g.tiles_n;
    }
    if constexpr ((MAP & 2) != 0) {
        if (tile_m >= g.cr_lo) tile_m += g.cr_n;
        if (tile_n >= g.cc_lo) tile_n += g.cc_n;
    }
    tile_m *= BM;
    tile_n *= BN;
    if (tile_m >= g.srlo && tile_m + BM <= g.srhi) return;
    if (tile_n >= g.sclo && tile_n + BN <= g.schi) return;
    if (tile_m >= g.sblo && tile_m + BM <= g.sbhi && tile_n >= g.sblo && tile_n + BN <= g.sbhi) return;
    if (g.ctr_wait) {
        if (threadIdx.x == 0) {
            unsigned it = 0;
            while (__hip_atomic_load(g.ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < g.ctr_wait) {
                if (++it >= kHandoverSpins) {
                    atomicOr(g.fail, 2);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        __syncthreads();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    // after the stores of a tile that another problem of the launch waits for
    unsigned* const hctr = g.ctr;
    const bool hcount = hctr && tile_n < g.ctr_chi && tile_n + BN > g.ctr_clo;
    auto hand_over = [hctr, hcount]() {
        if (hcount) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            __syncthreads();
            if (threadIdx.x == 0) __hip_atomic_fetch_add(hctr, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
    };
    // the tiles of a hand-over are the critical path of their launch: their
    // waves take issue priority over the trailing update's on shared SIMDs
    if (hcount || g.ctr_wait) __builtin_amdgcn_s_setprio(3);

    using V = typename Mfma<T>::v4;
    using VT = typename Vec<T>::type;
    constexpr int VW = Vec<T>::W;
    constexpr int NT = WM * WN * 64;
    constexpr int TM = BM / WM, TN = BN / WN;
    constexpr int FM = TM / 16, FN = TN / 16;
    constexpr int LA = BK + 4, LB = BN + 4;  // padded LDS rows (16-byte multiples)
    // K blocks of KB = 16 (MFMA steps s = 0..3 at lane group g take k = 4g + s)
    // or, for BK = 8, of KB = 8 (two steps, k = 2g + s)
    constexpr int KB = BK >= 16 ? 16 : 8, KS = KB / 4;
    static_assert(BK % KB == 0 && KS % VW == 0, "K chunk must be 8 or a multiple of 16");
    __shared__ __align__(16) T sA[2][2][BM][LA];
    __shared__ __align__(16) T sB[2][2][BK][LB];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave / WN, wn = wave % WN;
    const int l16 = lane & 15, g4 = lane >> 4;
    const T sgn = (T)g.alpha_re;  // +-1

    // M3 (three real products): rr = ArBr, ii = AiBi, ri = (Ar + Ai)(Br + Bi);
    // Re = rr - ii, Im = ri - rr - ii (C enters as rr = Cr, ri = Cr + Ci)
    static_assert(!M3 || TR, "3M products use the transposed accumulators");
    constexpr bool A3 = ACC3 || M3;
    V rr[FM][FN], ii[A3 ? FM : 1][A3 ? FN : 1], ri[FM][FN];
#pragma unroll
    for (int a = 0; a < FM; ++a)
#pragma unroll
        for (int b = 0; b < FN; ++b) {
            rr[a][b] = V{0, 0, 0, 0};
            ri[a][b] = V{0, 0, 0, 0};
            if constexpr (A3) ii[a][b] = V{0, 0, 0, 0};
        }

    // TR: the MFMA computes tiles of C^T (B as its first operand), so the
    // four accumulator elements of a lane are four consecutive columns of
    // one row of C, C[l16][4 g4 + r]: 16-byte loads and stores of C.  The B
    // operand column feeding MFMA row i is bperm(i) (FP64 accumulators hold
    // rows g4 + 4r, which the permutation sends to columns 4 g4 + r).
    auto bperm = [](int i) { return (TR && sizeof(T) == 8) ? 4 * (i & 3) + (i >> 2) : i; };
    // TR: the beta = 1 source is read after the last K chunk, into cx (the
    // other groups on the CU cover the load latency; reading it beside the
    // last chunk's MFMAs measured 1 % slower in the slice inverse), and added
    // in the epilogue; the accumulators start from zero, so the entries of C
    // are rounded once per launch instead of once per K step.
    constexpr bool LATE = TR;
    V cxr[LATE ? FM : 1][LATE ? FN : 1], cxi[LATE ? FM : 1][LATE ? FN : 1];
    // (the lambda reads the source through locals: kernel arguments taken
    // by reference inside a lambda end up in scratch memory)
    const T* const xsr = g.Xre;
    const T* const xsi = g.Xim;
    const int64_t ldxs = g.ldx;
    const int beta = g.beta;
    auto load_c = [&]() {
        const bool vx = (ldxs % 4) == 0 && (((uintptr_t)xsr | (uintptr_t)xsi) & 15) == 0;
#pragma unroll
        for (int a = 0; a < FM; ++a)
#pragma unroll
            for (int b = 0; b < FN; ++b) {
                cxr[a][b] = V{0, 0, 0, 0};
                cxi[a][b] = V{0, 0, 0, 0};
                const int gi = tile_m + wm * TM + a * 16 + l16;
                const int gj = tile_n + wn * TN + b * 16 + 4 * g4;
                if (gi < M) {
                    const int64_t o = (int64_t)gi * ldxs + gj;
                    if (vx && gj + 4 <= N) {
#pragma unroll
                        for (int u = 0; u < 4; u += VW) {
                            const VT xr = *(const VT*)(xsr + o + u);
                            const VT xi = *(const VT*)(xsi + o + u);
#pragma unroll
                            for (int v = 0; v < VW; ++v) {
                                cxr[a][b][u + v] = xr[v];
                                cxi[a][b][u + v] = xi[v];
                            }
                        }
                    } else {
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                            if (gj + r < N) {
                                cxr[a][b][r] = xsr[o + r];
                                cxi[a][b][r] = xsi[o + r];
                            }
                    }
                }
            }
    };
    if (g.beta) {
        if constexpr (!TR) {
#pragma unroll
            for (int a = 0; a < FM; ++a)
#pragma unroll
                for (int b = 0; b < FN; ++b) {
                    const int gj = tile_n + wn * TN + b * 16 + l16;
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int gi = tile_m + wm * TM + a * 16 + Mfma<T>::row(g4, r);
                        if (gi < M && gj < N) {
                            rr[a][b][r] = g.Xre[(int64_t)gi * g.ldx + gj];
                            ri[a][b][r] = g.Xim[(int64_t)gi * g.ldx + gj];
                        }
                    }
                }
        }
    }

    // global -> register staging of one K chunk; B column vectors are read up
    // to NVW, N rounded up to the vector width (padding inside the rows)
    const int NVW = (N + VW - 1) / VW * VW;
    constexpr int AV = BM * BK / VW, BV = BK * BN / VW;
    constexpr int AQ = (AV + NT - 1) / NT, BQ = (BV + NT - 1) / NT;
    VT va_r[AQ], va_i[AQ], vb_r[BQ], vb_i[BQ];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int q = 0; q < AQ; ++q) {
            const int e = tid + q * NT;
            if (AV % NT == 0 || e < AV) {
                const int i = e / (BK / VW), kv = (e % (BK / VW)) * VW;
                const int64_t o = (int64_t)min(tile_m + i, M - 1) * lda + k0 + kv;
                va_r[q] = *(const VT*)(Are + o);
                va_i[q] = *(const VT*)(Aim + o);
            }
        }
#pragma unroll
        for (int q = 0; q < BQ; ++q) {
            const int e = tid + q * NT;
            if (BV % NT == 0 || e < BV) {
                const int k = e / (BN / VW), nv = (e % (BN / VW)) * VW;
                const int64_t o = (int64_t)(k0 + k) * ldb + min(tile_n + nv, NVW - VW);
                vb_r[q] = *(const VT*)(Bre + o);
                vb_i[q] = *(const VT*)(Bim + o);
            }
        }
    };
    auto stash = [&](int buf) {
#pragma unroll
        for (int q = 0; q < AQ; ++q) {
            const int e = tid + q * NT;
            if (AV % NT == 0 || e < AV) {
                const int i = e / (BK / VW), kv = (e % (BK / VW)) * VW;
                *(VT*)&sA[buf][0][i][kv] = va_r[q] * sgn;
                *(VT*)&sA[buf][1][i][kv] = va_i[q] * sgn;
            }
        }
#pragma unroll
        for (int q = 0; q < BQ; ++q) {
            const int e = tid + q * NT;
            if (BV % NT == 0 || e < BV) {
                const int k = e / (BN / VW), nv = (e % (BN / VW)) * VW;
                *(VT*)&sB[buf][0][k][nv] = vb_r[q];
                *(VT*)&sB[buf][1][k][nv] = vb_i[q];
            }
        }
    };

    const int nk = K / BK;
    fetch(0);
    stash(0);
    __syncthreads();
    // MFMA work of the K chunk in LDS buffer buf
    auto chunk = [&](int buf) {
#pragma unroll
        for (int kb = 0; kb < BK; kb += KB) {
            T ar[FM][KS], ai[FM][KS], br[FN][KS], bi[FN][KS];
#pragma unroll
            for (int a = 0; a < FM; ++a) {
                const int row = wm * TM + a * 16 + l16;
#pragma unroll
                for (int u = 0; u < KS; u += VW) {
                    const VT xr = *(const VT*)&sA[buf][0][row][kb + KS * g4 + u];
                    const VT xi = *(const VT*)&sA[buf][1][row][kb + KS * g4 + u];
#pragma unroll
                    for (int v = 0; v < VW; ++v) {
                        ar[a][u + v] = xr[v];
                        ai[a][u + v] = xi[v];
                    }
                }
            }
#pragma unroll
            for (int b = 0; b < FN; ++b) {
                const int col = wn * TN + b * 16 + bperm(l16);
#pragma unroll
                for (int s2 = 0; s2 < KS; ++s2) {
                    br[b][s2] = sB[buf][0][kb + KS * g4 + s2][col];
                    bi[b][s2] = sB[buf][1][kb + KS * g4 + s2][col];
                }
            }
            // x * y accumulated into the tile of C (TR: of C^T)
            auto mm = [](T x, T y, V c) { return TR ? Mfma<T>::mma(y, x, c) : Mfma<T>::mma(x, y, c); };
            if constexpr (M3) {
#pragma unroll
                for (int s2 = 0; s2 < KS; ++s2) {
#pragma unroll
                    for (int a = 0; a < FM; ++a)
#pragma unroll
                        for (int b = 0; b < FN; ++b) rr[a][b] = mm(ar[a][s2], br[b][s2], rr[a][b]);
#pragma unroll
                    for (int a = 0; a < FM; ++a)
#pragma unroll
                        for (int b = 0; b < FN; ++b) ii[a][b] = mm(ai[a][s2], bi[b][s2], ii[a][b]);
                    T sa[FM], sb[FN];
#pragma unroll
                    for (int a = 0; a < FM; ++a) sa[a] = ar[a][s2] + ai[a][s2];
#pragma unroll
                    for (int b = 0; b < FN; ++b) sb[b] = br[b][s2] + bi[b][s2];
#pragma unroll
                    for (int a = 0; a < FM; ++a)
#pragma unroll
                        for (int b = 0; b < FN; ++b) ri[a][b] = mm(sa[a], sb[b], ri[a][b]);
                }
            } else
#pragma unroll
            for (int s2 = 0; s2 < KS; ++s2) {
#pragma unroll
                for (int a = 0; a < FM; ++a)
#pragma unroll
                    for (int b = 0; b < FN; ++b) rr[a][b] = mm(ar[a][s2], br[b][s2], rr[a][b]);
                if constexpr (ACC3) {
#pragma unroll
                    for (int a = 0; a < FM; ++a)
#pragma unroll
                        for (int b = 0; b < FN; ++b) ii[a][b] = mm(ai[a][s2], bi[b][s2], ii[a][b]);
                } else {
#pragma unroll
                    for (int a = 0; a < FM; ++a)
#pragma unroll
                        for (int b = 0; b < FN; ++b) rr[a][b] = mm(-ai[a][s2], bi[b][s2], rr[a][b]);
                }
#pragma unroll
                for (int a = 0; a < FM; ++a)
#pragma unroll
                    for (int b = 0; b < FN; ++b) ri[a][b] = mm(ar[a][s2], bi[b][s2], ri[a][b]);
#pragma unroll
                for (int a = 0; a < FM; ++a)
#pragma unroll
                    for (int b = 0; b < FN; ++b) ri[a][b] = mm(ai[a][s2], br[b][s2], ri[a][b]);
            }
        }
    };
    if constexpr (LATE) {
        // the last chunk peeled off: no prefetch there, the C reads after it
        for (int kc = 0; kc + 1 < nk; ++kc) {
            const int buf = kc & 1;
            fetch((kc + 1) * BK);
            __builtin_amdgcn_sched_barrier(0);  // keep the next chunk's loads ahead of the MFMAs
            chunk(buf);
            __builtin_amdgcn_sched_barrier(0);  // and their use after them
            stash(buf ^ 1);
            __syncthreads();
        }
        chunk((nk - 1) & 1);
        if (beta) load_c();
    } else {
        for (int kc = 0; kc < nk; ++kc) {
            const int buf = kc & 1;
            if (kc + 1 < nk) fetch((kc + 1) * BK);
            chunk(buf);
            if (kc + 1 < nk) stash(buf ^ 1);
            __syncthreads();
        }
    }

    // epilogue (the accumulators started from C when beta = 1)
    unsigned* const amax = AMAX ? g.amax : nullptr;
    float mloc = 0.f;
    if constexpr (TR) {
        const bool vc = (ldc % 4) == 0 && (((uintptr_t)Cre | (uintptr_t)Cim) & 15) == 0;
#pragma unroll
        for (int a = 0; a < FM; ++a)
#pragma unroll
            for (int b = 0; b < FN; ++b) {
                const int gi = tile_m + wm * TM + a * 16 + l16;
                const int gj = tile_n + wn * TN + b * 16 + 4 * g4;
                if (gi >= M || (gi >= g.srlo && gi < g.srhi)) continue;
                const bool rblk = gi >= g.sblo && gi < g.sbhi;
                T vr[4], vi[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    vr[r] = rr[a][b][r];
                    if constexpr (A3) vr[r] -= ii[a][b][r];
                    vi[r] = ri[a][b][r];
                    if constexpr (M3) vi[r] -= rr[a][b][r] + ii[a][b][r];
                    if (LATE && beta) {
                        vr[r] += cxr[a][b][r];
                        vi[r] += cxi[a][b][r];
                    }
                }
                const int64_t o = (int64_t)gi * ldc + gj;
                const bool whole = vc && gj + 4 <= N && !(gj + 4 > g.sclo && gj < g.schi) &&
                                   !(rblk && gj + 4 > g.sblo && gj < g.sbhi);
                if (whole) {
#pragma unroll
                    for (int u = 0; u < 4; u += VW) {
                        VT xr, xi;
#pragma unroll
                        for (int v = 0; v < VW; ++v) {
                            xr[v] = vr[u + v];
                            xi[v] = vi[u + v];
                        }
                        *(VT*)(Cre + o + u) = xr;
                        *(VT*)(Cim + o + u) = xi;
                    }
                    if (AMAX && amax) {
#pragma unroll
                        for (int r = 0; r < 4; ++r) mloc = fmaxf(mloc, (float)(vr[r] * vr[r] + vi[r] * vi[r]));
                    }
                } else {
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int c = gj + r;
                        if (c < N && !(c >= g.sclo && c < g.schi) && !(rblk && c >= g.sblo && c < g.sbhi)) {
                            Cre[o + r] = vr[r];
                            Cim[o + r] = vi[r];
                            if (AMAX && amax) mloc = fmaxf(mloc, (float)(vr[r] * vr[r] + vi[r] * vi[r]));
                        }
                    }
                }
                if (gi >= g.y_rlo && gi < g.y_rhi) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int c = gj + r;
                        if (c < N && c >= g.y_clo && c < g.y_chi && !(c >= g.sclo && c < g.schi) &&
                            !(rblk && c >= g.sblo && c < g.sbhi)) {
                            const int64_t oy = (int64_t)(gi - g.y_rlo) * g.ldy + (c - g.y_clo);
                            g.Yre[oy] = vr[r];
                            g.Yim[oy] = vi[r];
                        }
                    }
                }
            }
        if (AMAX && amax) wave_amax(amax, mloc);
        hand_over();
        return;
    }
#pragma unroll
    for (int a = 0; a < FM; ++a)
#pragma unroll
        for (int b = 0; b < FN; ++b) {
            const int gj = tile_n + wn * TN + b * 16 + l16;
            const bool cok = gj < N && !(gj >= g.sclo && gj < g.schi);
            const bool cblk = gj >= g.sblo && gj < g.sbhi;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gi = tile_m + wm * TM + a * 16 + Mfma<T>::row(g4, r);
                if (cok && gi < M && !(gi >= g.srlo && gi < g.srhi) && !(cblk && gi >= g.sblo && gi < g.sbhi)) {
                    T vr = rr[a][b][r];
                    if constexpr (ACC3) vr -= ii[a][b][r];
                    Cre[(int64_t)gi * ldc + gj] = vr;
                    Cim[(int64_t)gi * ldc + gj] = ri[a][b][r];
                    if (AMAX && amax) mloc = fmaxf(mloc, (float)(vr * vr + ri[a][b][r] * ri[a][b][r]));
                    if (gi >= g.y_rlo && gi < g.y_rhi && gj >= g.y_clo && gj < g.y_chi) {
                        const int64_t oy = (int64_t)(gi - g.y_rlo) * g.ldy + (gj - g.y_clo);
                        g.Yre[oy] = vr;
                        g.Yim[oy] = ri[a][b][r];
                    }
                }
            }
        }
    if (AMAX && amax) wave_amax(amax, mloc);
    hand_over();
}

// FP32 64x64 tiles: four work groups per CU (four waves per SIMD, at most
// 128 registers); the three-product accumulators need all of them.  FP64
// 64x64 tiles with 8-deep K chunks: three (at most 168 registers).
template <typename T, int WM, int WN, int BK>
constexpr int gemm_min_blocks() {
    return (sizeof(T) == 4 && WM * WN == 4) ? 4 : (WM * WN == 4 && BK == 8) ? 3 : 1;
}

template <typename T, int BM, int BN, int BK, int WM, int WN, bool ACC3, int MAP, bool TR, bool M3, bool AMAX = false>
__global__ __launch_bounds__(WM* WN * 64, (gemm_min_blocks<T, WM, WN, BK>())) void cgemm2_kernel(
    GemmArgs<T> g0, GemmArgs<T> g1, GemmArgs<T> g2, int total) {
    int bid = (int)blockIdx.x;
    if constexpr ((MAP & 1) != 0) {
        const int q = total >> 3, rem = total & 7, x = bid & 7, loc = bid >> 3;
        bid = x * q + min(x, rem) + loc;
    }
    // one inlined copy of the tile per problem: each reads its own arguments
    // at fixed offsets instead of selecting among the three per field
    if (bid < g0.tiles) {
        cgemm2_tile<T, BM, BN, BK, WM, WN, ACC3, MAP, TR, M3, AMAX>(g0, bid);
        return;
    }
    bid -= g0.tiles;
    if (bid < g1.tiles) {
        cgemm2_tile<T, BM, BN, BK, WM, WN, ACC3, MAP, TR, M3, AMAX>(g1, bid);
        return;
    }
    cgemm2_tile<T, BM, BN, BK, WM, WN, ACC3, MAP, TR, M3, AMAX>(g2, bid - g1.tiles);
}

// Lean single-problem tile kernel: C = alpha A B (+ C if BETA) with alpha =
// -1 or +1 (folded into A), C itself the beta = 1 source, three (M3, FP32)
// or four real products, 64 x 64 tiles over the (optionally compressed)
// tile grid of set_tiles, one tile per work group; SPLIT: chunk z of an
// nz-way split along K writes its partial to C + z * zstride.  The
// arithmetic is that of cgemm2_tile<T, 64, 64, BK, 2, 2, false, MAP, true,
// M3> (the same results, bit for bit); left out is what these products do
// not use (further problems in the launch, second destination, hand-over,
// multiplier maximum), and C is read after the last K chunk rather than
// beside it, the other groups on the CU covering those reads.  The slice
// inverse's bulk update at N = 4096 takes, alone, 111 us against 121 (FP32,
// b = 128, three products), 75 against 87 (FP32, b = 64) and 162 against
// 176 (FP64, b = 64).  (The tile loop lets a group take several tiles;
// launched with one group per tile it keeps the one-tile groups' turnover
// of CU slots, which the launches beside the bulk update need.)
template <typename T, bool M3, bool BETA, bool SPLIT>
__global__ __launch_bounds__(256, sizeof(T) == 8 ? 3 : 4) void lean_gemm_kernel(GemmArgs<T> g) {
    using V = typename Mfma<T>::v4;
    using VT = typename Vec<T>::type;
    // K chunks of 16 (FP32) or 8 (FP64: 42 KB of LDS, 168 registers, three
    // groups per CU); MFMA step k at lane group g4 takes chunk row KS g4 + k
    constexpr int BM = 64, BN = 64, BK = sizeof(T) == 8 ? 8 : 16, WN = 2;
    constexpr int VW = Vec<T>::W;
    constexpr int TM = 32, TN = 32, FM = 2, FN = 2, KS = BK / 4;
    constexpr int LA = BK + 4, LB = BN + 4;  // padded LDS rows (16-byte multiples)
    static_assert(BM * BK / VW == 256 && BK * BN / VW == 256, "one vector per thread and operand");
    static_assert(KS == VW, "a lane's A values of a chunk are one vector");
    static_assert(!M3 || sizeof(T) == 4, "three real products in FP32 only");
    // B column feeding MFMA row i: FP64 accumulators hold rows g4 + 4r, which
    // this permutation sends to columns 4 g4 + r
    auto bperm = [](int i) { return sizeof(T) == 8 ? 4 * (i & 3) + (i >> 2) : i; };
    __shared__ __align__(16) T sA[2][2][BM][LA];
    __shared__ __align__(16) T sB[2][2][BK][LB];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave / WN, wn = wave % WN;
    const int l16 = lane & 15, g4 = lane >> 4;
    const T sgn = (T)g.alpha_re;
    const int M = g.M, N = g.N;
    const int64_t lda = g.lda, ldb = g.ldb, ldc = g.ldc;
    const int NVW = (N + VW - 1) / VW * VW;
    const int tiles_n = g.tiles_n, ntiles = g.tiles;
    const int per = SPLIT ? ntiles / g.nz : ntiles;
    const int cr_lo = g.cr_lo, cr_n = g.cr_n, cc_lo = g.cc_lo, cc_n = g.cc_n;
    const int srlo = g.srlo, srhi = g.srhi, sclo = g.sclo, schi = g.schi;
    const int sblo = g.sblo, sbhi = g.sbhi;
    const bool vc = (ldc % 4) == 0 && (((uintptr_t)g.Cre | (uintptr_t)g.Cim) & 15) == 0 && (g.zstride % 4) == 0;
    // the thread's operand vectors: A row ai_, columns akv; B row bk, columns bnv
    const int ai_ = tid / (BK / VW), akv = (tid % (BK / VW)) * VW;
    const int bk = tid / (BN / VW), bnv = (tid % (BN / VW)) * VW;
    for (int t = (int)blockIdx.x; t < ntiles; t += (int)gridDim.x) {
        // chunk z of the split: K range [z kc, z kc + kc) (the last one shorter)
        const int z = SPLIT ? t / per : 0, lin = t - z * per;
        const int64_t koff = (int64_t)z * g.kc;
        const int nk = (SPLIT ? (int)min((int64_t)g.kc, (int64_t)g.K - koff) : g.K) / BK;
        const T* __restrict__ Are = g.Are + koff;
        const T* __restrict__ Aim = g.Aim + koff;
        const T* __restrict__ Bre = g.Bre + koff * ldb;
        const T* __restrict__ Bim = g.Bim + koff * ldb;
        T* const Cre = g.Cre + z * g.zstride;
        T* const Cim = g.Cim + z * g.zstride;
        int tr = lin / tiles_n, tc = lin - tr * tiles_n;
        if (tr >= cr_lo) tr += cr_n;
        if (tc >= cc_lo) tc += cc_n;
        const int tm = tr * BM, tn = tc * BN;
        if ((tm >= srlo && tm + BM <= srhi) || (tn >= sclo && tn + BN <= schi) ||
            (tm >= sblo && tm + BM <= sbhi && tn >= sblo && tn + BN <= sbhi))
            continue;
        VT va_r, va_i, vb_r, vb_i;
        auto fetch = [&](int k0) {
            const int64_t oa = (int64_t)min(tm + ai_, M - 1) * lda + k0 + akv;
            va_r = *(const VT*)(Are + oa);
            va_i = *(const VT*)(Aim + oa);
            const int64_t ob = (int64_t)(k0 + bk) * ldb + min(tn + bnv, NVW - VW);
            vb_r = *(const VT*)(Bre + ob);
            vb_i = *(const VT*)(Bim + ob);
        };
        auto stash = [&](int buf) {
            *(VT*)&sA[buf][0][ai_][akv] = va_r * sgn;
            *(VT*)&sA[buf][1][ai_][akv] = va_i * sgn;
            *(VT*)&sB[buf][0][bk][bnv] = vb_r;
            *(VT*)&sB[buf][1][bk][bnv] = vb_i;
        };
        // M3: rr = ArBr, ii = AiBi, ri = (Ar + Ai)(Br + Bi); else rr = Re, ri = Im
        V rr[FM][FN], ii[M3 ? FM : 1][M3 ? FN : 1], ri[FM][FN];
#pragma unroll
        for (int x = 0; x < FM; ++x)
#pragma unroll
            for (int y = 0; y < FN; ++y) {
                rr[x][y] = V{0, 0, 0, 0};
                ri[x][y] = V{0, 0, 0, 0};
                if constexpr (M3) ii[x][y] = V{0, 0, 0, 0};
            }
        // one 16-deep K chunk from LDS buffer buf (MFMA step k at lane group
        // g4 takes row 4 g4 + k of the chunk; the MFMA computes tiles of C^T)
        auto chunk = [&](int buf) {
            T ar[FM][KS], ai[FM][KS], br[FN][KS], bi[FN][KS];
#pragma unroll
            for (int x = 0; x < FM; ++x) {
                const int row = wm * TM + x * 16 + l16;
                const VT xr = *(const VT*)&sA[buf][0][row][KS * g4];
                const VT xi = *(const VT*)&sA[buf][1][row][KS * g4];
#pragma unroll
                for (int v = 0; v < VW; ++v) {
                    ar[x][v] = xr[v];
                    ai[x][v] = xi[v];
                }
            }
#pragma unroll
            for (int y = 0; y < FN; ++y) {
                const int col = wn * TN + y * 16 + bperm(l16);
#pragma unroll
                for (int k = 0; k < KS; ++k) {
                    br[y][k] = sB[buf][0][KS * g4 + k][col];
                    bi[y][k] = sB[buf][1][KS * g4 + k][col];
                }
            }
            auto mm = [](T a, T b, V c) { return Mfma<T>::mma(b, a, c); };
#pragma unroll
            for (int k = 0; k < KS; ++k) {
                if constexpr (M3) {
#pragma unroll
                    for (int x = 0; x < FM; ++x)
#pragma unroll
                        for (int y = 0; y < FN; ++y) rr[x][y] = mm(ar[x][k], br[y][k], rr[x][y]);
#pragma unroll
                    for (int x = 0; x < FM; ++x)
#pragma unroll
                        for (int y = 0; y < FN; ++y) ii[x][y] = mm(ai[x][k], bi[y][k], ii[x][y]);
                    T sa[FM], sb[FN];
#pragma unroll
                    for (int x = 0; x < FM; ++x) sa[x] = ar[x][k] + ai[x][k];
#pragma unroll
                    for (int y = 0; y < FN; ++y) sb[y] = br[y][k] + bi[y][k];
#pragma unroll
                    for (int x = 0; x < FM; ++x)
#pragma unroll
                        for (int y = 0; y < FN; ++y) ri[x][y] = mm(sa[x], sb[y], ri[x][y]);
                } else {
#pragma unroll
                    for (int x = 0; x < FM; ++x)
#pragma unroll
                        for (int y = 0; y < FN; ++y) rr[x][y] = mm(ar[x][k], br[y][k], rr[x][y]);
#pragma unroll
                    for (int x = 0; x < FM; ++x)
#pragma unroll
                        for (int y = 0; y < FN; ++y) rr[x][y] = mm(-ai[x][k], bi[y][k], rr[x][y]);
#pragma unroll
                    for (int x = 0; x < FM; ++x)
#pragma unroll
                        for (int y = 0; y < FN; ++y) ri[x][y] = mm(ar[x][k], bi[y][k], ri[x][y]);
#pragma unroll
                    for (int x = 0; x < FM; ++x)
#pragma unroll
                        for (int y = 0; y < FN; ++y) ri[x][y] = mm(ai[x][k], br[y][k], ri[x][y]);
                }
            }
        };
        __syncthreads();  // every wave is done with the previous tile's buffers
        fetch(0);
        stash(0);
        __syncthreads();
        for (int kc = 0; kc + 1 < nk; ++kc) {
            const int buf = kc & 1;
            fetch((kc + 1) * BK);
            __builtin_amdgcn_sched_barrier(0);  // the next chunk's loads ahead of the MFMAs
            chunk(buf);
            __builtin_amdgcn_sched_barrier(0);  // and their use after them
            stash(buf ^ 1);
            __syncthreads();
        }
        chunk((nk - 1) & 1);
        // C += the tile (a lane holds C[row][4 g4 + r], r = 0..3), masked to
        // the rows and columns outside the skips; all of the lane's C reads
        // are issued before the first store
        V cxr[BETA ? FM : 1][BETA ? FN : 1], cxi[BETA ? FM : 1][BETA ? FN : 1];
        if constexpr (BETA)
#pragma unroll
        for (int x = 0; x < FM; ++x)
#pragma unroll
            for (int y = 0; y < FN; ++y) {
                cxr[x][y] = V{0, 0, 0, 0};
                cxi[x][y] = V{0, 0, 0, 0};
                const int gi = tm + wm * TM + x * 16 + l16;
                const int gj = tn + wn * TN + y * 16 + 4 * g4;
                if (gi >= M) continue;
                const int64_t o = (int64_t)gi * ldc + gj;
                if (vc && gj + 4 <= N) {
#pragma unroll
                    for (int u = 0; u < 4; u += VW) {
                        const VT xr = *(const VT*)(Cre + o + u);
                        const VT xi = *(const VT*)(Cim + o + u);
#pragma unroll
                        for (int v = 0; v < VW; ++v) {
                            cxr[x][y][u + v] = xr[v];
                            cxi[x][y][u + v] = xi[v];
                        }
                    }
                } else {
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        if (gj + r < N) {
                            cxr[x][y][r] = Cre[o + r];
                            cxi[x][y][r] = Cim[o + r];
                        }
                }
            }
#pragma unroll
        for (int x = 0; x < FM; ++x)
#pragma unroll
            for (int y = 0; y < FN; ++y) {
                const int gi = tm + wm * TM + x * 16 + l16;
                const int gj = tn + wn * TN + y * 16 + 4 * g4;
                if (gi >= M || (gi >= srlo && gi < srhi)) continue;
                const bool rblk = gi >= sblo && gi < sbhi;
                const int64_t o = (int64_t)gi * ldc + gj;
                const bool whole = vc && gj + 4 <= N && !(gj + 4 > sclo && gj < schi) &&
                                   !(rblk && gj + 4 > sblo && gj < sbhi);
                T vr[4], vi[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    vr[r] = rr[x][y][r];
                    vi[r] = ri[x][y][r];
                    if constexpr (M3) {
                        vr[r] -= ii[x][y][r];
                        vi[r] -= rr[x][y][r] + ii[x][y][r];
                    }
                    if constexpr (BETA) {
                        vr[r] += cxr[x][y][r];
                        vi[r] += cxi[x][y][r];
                    }
                }
                if (whole) {
#pragma unroll
                    for (int u = 0; u < 4; u += VW) {
                        VT xr, xi;
#pragma unroll
                        for (int v = 0; v < VW; ++v) {
                            xr[v] = vr[u + v];
                            xi[v] = vi[u + v];
                        }
                        *(VT*)(Cre + o + u) = xr;
                        *(VT*)(Cim + o + u) = xi;
                    }
                } else {
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int c = gj + r;
                        if (c < N && !(c >= sclo && c < schi) && !(rblk && c >= sblo && c < sbhi)) {
                            Cre[o + r] = vr[r];
                            Cim[o + r] = vi[r];
                        }
                    }
                }
            }
    }
}

template <typename T>
struct GemmCfg;
// FP64: 64x64 tiles, 4 waves (32x32 per wave); 128x128, 8 waves (64x32).
// FP32: same shapes with a deeper K chunk.
// BK2: K chunk of the streamlined 64x64 kernel.  FP64 takes 8: 42 KB of LDS
// and at most 168 registers, three work groups per CU instead of two (the
// rank-64 trailing update at N = 4096: 172 instead of 198 us alone, the FP64
// slice inverse 13.0 instead of 14.8 ms, with the panel and next-block
// launches on the same footprint so that they take the places the update's
// groups free).
template <>
struct GemmCfg<double> {
    static constexpr int BK_SMALL = 16, BK_BIG = 8, BK2 = 8;
};
template <>
struct GemmCfg<float> {
    static constexpr int BK_SMALL = 32, BK_BIG = 32, BK2 = 16;
};

template <typename T>
static GemmArgs<T> gemm_args(int64_t M, int64_t N, int64_t K, CMat<T> A, CMat<T> B, CMat<T> C, cd alpha, int beta,
                             Skip sk) {
    // vector loads need 16-byte aligned rows in A and B
    const int VW = 16 / (int)sizeof(T);
    GemmArgs<T> g;
    g.Are = A.re;
    g.Aim = A.im;
    g.lda = A.ld;
    g.Bre = B.re;
    g.Bim = B.im;
    g.ldb = B.ld;
    g.Cre = C.re;
    g.Cim = C.im;
    g.ldc = C.ld;
    g.Xre = C.re;
    g.Xim = C.im;
    g.ldx = C.ld;
    g.Yre = nullptr;
    g.Yim = nullptr;
    g.ldy = 0;
    g.y_rlo = g.y_rhi = g.y_clo = g.y_chi = 0;
    g.M = (int)M;
    g.N = (int)N;
    g.K = (int)K;
    g.alpha_re = alpha.real();
    g.alpha_im = alpha.imag();
    g.beta = beta;
    g.srlo = (int)sk.rlo;
    g.srhi = (int)sk.rhi;
    g.sclo = (int)sk.clo;
    g.schi = (int)sk.chi;
    g.sblo = (int)sk.blo;
    g.sbhi = (int)sk.bhi;
    g.nz = 1;
    g.kc = (int)K;
    g.zstride = 0;
    g.cr_lo = g.cc_lo = INT32_MAX;
    g.cr_n = g.cc_n = 0;
    g.vec_ok = (A.ld % VW == 0) && (B.ld % VW == 0) && ((uintptr_t)A.re % 16 == 0) && ((uintptr_t)A.im % 16 == 0) &&
               ((uintptr_t)B.re % 16 == 0) && ((uintptr_t)B.im % 16 == 0);
    g.amax = nullptr;
    g.ctr = nullptr;
    g.ctr_clo = g.ctr_chi = 0;
    g.ctr_wait = 0;
    g.fail = nullptr;
    return g;
}

template <typename T, int BM, int BN, int BK, int WM, int WN>
static void launch_pair(hipStream_t s, const GemmArgs<T>& g0_, const GemmArgs<T>* g1_, int use3m) {
    GemmArgs<T> g0 = g0_, g1 = g1_ ? *g1_ : g0_;
    auto tiles = [](GemmArgs<T>& g) {
        g.tiles_n = (g.N + BN - 1) / BN;
        g.tiles = g.tiles_n * ((g.M + BM - 1) / BM) * g.nz;
        if (g.M <= 0 || g.N <= 0) g.tiles = 0;
    };
    tiles(g0);
    tiles(g1);
    const int64_t nblocks = (int64_t)g0.tiles + (g1_ ? g1.tiles : 0);
    if (nblocks <= 0) return;
    dim3 grid((unsigned)nblocks);
    dim3 block(WM * WN * 64);
    if (use3m)
        hipLaunchKernelGGL((cgemm_kernel<T, BM, BN, BK, WM, WN, true>), grid, block, 0, s, g0, g1);
    else
        hipLaunchKernelGGL((cgemm_kernel<T, BM, BN, BK, WM, WN, false>), grid, block, 0, s, g0, g1);
}

// Tile grid of a problem (tiles_n, tiles).  COMPRESS: whole skipped tile rows
// and columns left out of the grid (a skip that reaches past the last row or
// column covers the tiles up to the end of the grid).
template <typename T, int BM, int BN, bool COMPRESS>
static void set_tiles(GemmArgs<T>& g) {
    int tm = (g.M + BM - 1) / BM, tn = (g.N + BN - 1) / BN;
    if constexpr (COMPRESS) {
        const int rhi = g.srhi >= g.M ? tm * BM : g.srhi, chi = g.schi >= g.N ? tn * BN : g.schi;
        if (rhi > g.srlo && g.srlo % BM == 0 && rhi % BM == 0) {
            g.cr_lo = g.srlo / BM;
            g.cr_n = (rhi - g.srlo) / BM;
            tm -= g.cr_n;
        }
        if (chi > g.sclo && g.sclo % BN == 0 && chi % BN == 0) {
            g.cc_lo = g.sclo / BN;
            g.cc_n = (chi - g.sclo) / BN;
            tn -= g.cc_n;
        }
    }
    g.tiles_n = tn;
    g.tiles = tn * tm * g.nz;
}

// cgemm2_kernel launch for one or two problems.  Returns false (launching
// nothing) when a problem is outside the kernel's preconditions.
template <typename T, int BM, int BN, int BK, int WM, int WN, bool ACC3 = sizeof(T) == 4, int MAP = 0,
          bool TR = true, bool M3 = false>
static bool launch_pair2(hipStream_t s, const GemmArgs<T>& g0_, const GemmArgs<T>* g1_,
                         const GemmArgs<T>* g2_ = nullptr, hipEvent_t stop = nullptr) {
    constexpr int VW = Vec<T>::W;
    auto fits = [](const GemmArgs<T>& g) {
        const bool unit = g.alpha_im == 0.0 && (g.alpha_re == 1.0 || g.alpha_re == -1.0);
        const bool kok = g.K > 0 && g.K % BK == 0 && (g.nz <= 1 || g.kc % BK == 0);
        // B column vectors may reach past N up to the next multiple of VW,
        // which must stay inside the row (ldb)
        const int64_t nv = ((int64_t)g.N + VW - 1) / VW * VW;
        return unit && kok && g.vec_ok && nv <= g.ldb && g.M > 0;
    };
    if (!fits(g0_) || (g1_ && !fits(*g1_)) || (g2_ && !fits(*g2_))) return false;
    GemmArgs<T> g0 = g0_, g1 = g1_ ? *g1_ : g0_, g2 = g2_ ? *g2_ : g0_;
    auto tiles = [](GemmArgs<T>& g) { set_tiles<T, BM, BN, (MAP & 2) != 0>(g); };
    tiles(g0);
    tiles(g1);
    tiles(g2);
    if (!g1_) g1.tiles = 0;
    if (!g2_) g2.tiles = 0;
    const int64_t nblocks = (int64_t)g0.tiles + g1.tiles + g2.tiles;
    if (nblocks <= 0) {
        if (stop) RGF_HIP(hipEventRecord(stop, s));
        return true;
    }
    const bool am = g0.amax || (g1_ && g1.amax) || (g2_ && g2.amax);
    auto kern = am ? cgemm2_kernel<T, BM, BN, BK, WM, WN, ACC3, MAP, TR, M3, true>
                   : cgemm2_kernel<T, BM, BN, BK, WM, WN, ACC3, MAP, TR, M3, false>;
    if (stop) {
        // the event completes with the kernel (no separate marker packet)
        hipExtLaunchKernelGGL(kern, dim3((unsigned)nblocks), dim3(WM * WN * 64), 0, s, nullptr, stop, 0, g0, g1, g2,
                              (int)nblocks);
    } else {
        hipLaunchKernelGGL(kern, dim3((unsigned)nblocks), dim3(WM * WN * 64), 0, s, g0, g1, g2, (int)nblocks);
    }
    return true;
}

// launch_pair2 with four (use3m = 0) or three (use3m != 0) real products
template <typename T, int BM, int BN, int BK, int WM, int WN>
static bool launch_cplx(int use3m, hipStream_t s, const GemmArgs<T>& g0, const GemmArgs<T>* g1 = nullptr,
                        const GemmArgs<T>* g2 = nullptr, hipEvent_t stop = nullptr) {
    if (use3m) return launch_pair2<T, BM, BN, BK, WM, WN, false, 0, true, true>(s, g0, g1, g2, stop);
    return launch_pair2<T, BM, BN, BK, WM, WN, false>(s, g0, g1, g2, stop);
}

// One problem on lean_gemm_kernel, one work group per tile; COMPRESS: whole
// skipped tile rows and columns left out of the grid.  Returns false
// (launching nothing) outside its preconditions: alpha = +-1, beta = 0 or 1
// from C itself, no second destination, maximum or hand-over, K (and the
// split's chunk) a multiple of the K chunk (16 FP32, 8 FP64), three
// products in FP32 only, 16-byte aligned operand rows.
template <typename T, bool COMPRESS = false>
static bool launch_lean(hipStream_t s, const GemmArgs<T>& g_, int use3m, hipEvent_t stop = nullptr) {
    constexpr int VW = Vec<T>::W, BK = sizeof(T) == 8 ? 8 : 16;
    const GemmArgs<T>& c = g_;
    const bool unit = c.alpha_im == 0.0 && (c.alpha_re == 1.0 || c.alpha_re == -1.0);
    const int64_t nv = ((int64_t)c.N + VW - 1) / VW * VW;
    const bool ysec = c.Yre && c.y_rhi > c.y_rlo && c.y_chi > c.y_clo;
    const bool src_c = c.Xre == c.Cre && c.Xim == c.Cim && c.ldx == c.ldc;
    const bool split = c.nz > 1;
    if (!unit || (c.beta != 0 && c.beta != 1) || (c.beta && (!src_c || split)) || c.K <= 0 || c.K % BK != 0 ||
        (split && c.kc % BK != 0) || !c.vec_ok || nv > c.ldb || c.M <= 0 || c.N <= 0 || c.amax || c.ctr ||
        c.ctr_wait || ysec || (use3m && sizeof(T) != 4))
        return false;
    GemmArgs<T> g = g_;
    set_tiles<T, 64, 64, COMPRESS>(g);
    if (g.tiles <= 0) {
        if (stop) RGF_HIP(hipEventRecord(stop, s));
        return true;
    }
    auto go = [&](auto kern) {
        if (stop)
            hipExtLaunchKernelGGL(kern, dim3(g.tiles), dim3(256), 0, s, nullptr, stop, 0, g);
        else
            hipLaunchKernelGGL(kern, dim3(g.tiles), dim3(256), 0, s, g);
    };
    if constexpr (sizeof(T) == 4) {
        if (use3m) {
            if (split)
                go(lean_gemm_kernel<T, true, false, true>);
            else if (c.beta)
                go(lean_gemm_kernel<T, true, true, false>);
            else
                go(lean_gemm_kernel<T, true, false, false>);
            return true;
        }
    }
    if (split)
        go(lean_gemm_kernel<T, false, false, true>);
    else if (c.beta)
        go(lean_gemm_kernel<T, false, true, false>);
    else
        go(lean_gemm_kernel<T, false, false, false>);
    return true;
}

template <typename T, int BM, int BN, int BK, int WM, int WN>
static void launch_cgemm(hipStream_t s, int64_t M, int64_t N, int64_t K, CMat<T> A, CMat<T> B, CMat<T> C,
                         cd alpha, int beta, Skip sk, int use3m) {
    const GemmArgs<T> g = gemm_args<T>(M, N, K, A, B, C, alpha, beta, sk);
    launch_pair<T, BM, BN, BK, WM, WN>(s, g, nullptr, use3m);
}

// K chunks for a split-K product with `tiles` 64x64 output tiles: about
// 1536 work groups in one launch (measured better than one or two full
// rounds of resident groups), chunks of at least 256 for K >= 1024 and of
// at least 64 below (thin products of small slices: without the split their
// few tiles run on a few CUs); 0 = no split.
template <typename T>
static int64_t splitk_chunks(int64_t tiles, int64_t K, size_t ws_elems, int64_t per, int64_t& kc_out) {
    constexpr int BK = GemmCfg<T>::BK_SMALL;
    int64_t nz = std::min<int64_t>((1536 + tiles - 1) / tiles, K / (K >= 1024 ? 256 : 64));
    nz = std::min<int64_t>(nz, (int64_t)(ws_elems / per));
    if (nz < 2) return 0;
    kc_out = ((K + nz - 1) / nz + BK - 1) / BK * BK;
    nz = (K + kc_out - 1) / kc_out;
    return nz >= 2 ? nz : 0;
}

// C = alpha * sum_z P_z + beta * C over the split-K partial products
template <typename T>
__global__ void splitk_reduce_kernel(int64_t M, int64_t N, int nz, const T* P, int64_t pstride, T* Cre, T* Cim,
                                     int64_t ldc, double ar, double ai, int beta) {
    const int64_t total = M * N;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        double sr = 0, si = 0;
        for (int z = 0; z < nz; ++z) {
            sr += (double)P[z * pstride + e];
            si += (double)P[z * pstride + total + e];
        }
        const int64_t i = e / N, j = e % N;
        const int64_t o = i * ldc + j;
        double vr = ar * sr - ai * si, vi = ar * si + ai * sr;
        if (beta) {
            vr += (double)Cre[o];
            vi += (double)Cim[o];
        }
        Cre[o] = (T)vr;
        Cim[o] = (T)vi;
    }
}

template <typename T>
void cgemm_skip(hipStream_t s, int64_t M, int64_t N, int64_t K, CMat<T> A, CMat<T> B, CMat<T> C, cd alpha,
                int beta, Skip sk, int use3m, T* ws, size_t ws_elems) {
    if (M <= 0 || N <= 0) return;
    const bool any_skip = sk.rhi > sk.rlo || sk.chi > sk.clo || sk.bhi > sk.blo;
    // split-K for thin products with a long inner dimension
    if (ws && !any_skip && K >= 256) {
        const int64_t tiles = ((M + 63) / 64) * ((N + 63) / 64);
        if (tiles < 512) {
            const int64_t per = 2 * M * N;  // elements of one partial (two planes)
            int64_t kc = 0;
            const int64_t nz = splitk_chunks<T>(tiles, K, ws_elems, per, kc);
            if (nz >= 2) {
                // all chunks in one launch; chunk z writes partial z of the workspace
                GemmArgs<T> g = gemm_args<T>(M, N, K, A, B, CMat<T>(ws, ws + M * N, N), cd(1, 0), 0, Skip{});
                g.nz = (int)nz;
                g.kc = (int)kc;
                g.zstride = per;
                if (!launch_lean<T>(s, g, use3m) && !launch_cplx<T, 64, 64, GemmCfg<T>::BK2, 2, 2>(use3m, s, g))
                    launch_pair<T, 64, 64, GemmCfg<T>::BK_SMALL, 2, 2>(s, g, nullptr, use3m);
                hipLaunchKernelGGL((splitk_reduce_kernel<T>), dim3(grid_for(M * N)), dim3(256), 0, s, M, N, (int)nz,
                                   (const T*)ws, per, C.re, C.im, C.ld, alpha.real(), alpha.imag(), beta);
                RGF_HIP(hipGetLastError());
                return;
            }
        }
    }
    if (K <= 0) {
        if (!beta) fill_zero<T>(s, M, N, C);
        return;
    }
    if (M > INT32_MAX || N > INT32_MAX || K > INT32_MAX) throw Error(RGF_ERR_INVALID_ARGUMENT, "GEMM too large");
    const int64_t tiles_big = ((M + 127) / 128) * ((N + 127) / 128);
    const int64_t tiles_mid = ((M + 63) / 64) * ((N + 63) / 64);
    {
        // streamlined kernel: 64x64 tiles (4 waves, K chunk 16) when there
        // are enough tiles, else 32x32
        const GemmArgs<T> g = gemm_args<T>(M, N, K, A, B, C, alpha, beta, sk);
        const bool done = tiles_mid >= 256 ? (launch_lean<T>(s, g, use3m) ||
                                              launch_cplx<T, 64, 64, GemmCfg<T>::BK2, 2, 2>(use3m, s, g))
                                           : launch_cplx<T, 32, 32, 16, 1, 1>(use3m, s, g);
        if (done) {
            RGF_HIP(hipGetLastError());
            return;
        }
    }
    // FP64 runs best on 128x128 tiles (8 waves); FP32 on 64x64 tiles, which
    // keep two or three work groups per CU
    if (sizeof(T) == 8 && tiles_big >= 1024)
        launch_cgemm<T, 128, 128, GemmCfg<T>::BK_BIG, 2, 4>(s, M, N, K, A, B, C, alpha, beta, sk, use3m);
    else if (tiles_mid >= 512)
        launch_cgemm<T, 64, 64, GemmCfg<T>::BK_SMALL, 2, 2>(s, M, N, K, A, B, C, alpha, beta, sk, use3m);
    else
        launch_cgemm<T, 32, 32, GemmCfg<T>::BK_SMALL, 1, 1>(s, M, N, K, A, B, C, alpha, beta, sk, use3m);
    RGF_HIP(hipGetLastError());
}

// Two small products in one launch; `stop`, if given, is recorded when the
// launch completes.
template <typename T>
static void cgemm_pair(hipStream_t s, const GemmArgs<T>& g0, const GemmArgs<T>& g1, int use3m,
                       hipEvent_t stop = nullptr) {
    if (!launch_cplx<T, 64, 64, GemmCfg<T>::BK2, 2, 2>(use3m, s, g0, &g1, nullptr, stop)) {
        launch_pair<T, 32, 32, GemmCfg<T>::BK_SMALL, 1, 1>(s, g0, &g1, use3m);
        if (stop) RGF_HIP(hipEventRecord(stop, s));
    }
    RGF_HIP(hipGetLastError());
}

// Three independent products in one launch (two launches on the fallback).
template <typename T>
static void cgemm_triple(hipStream_t s, const GemmArgs<T>& g0, const GemmArgs<T>& g1, const GemmArgs<T>& g2,
                         int use3m, hipEvent_t stop = nullptr) {
    if (!launch_cplx<T, 64, 64, GemmCfg<T>::BK2, 2, 2>(use3m, s, g0, &g1, &g2, stop)) {
        launch_pair<T, 32, 32, GemmCfg<T>::BK_SMALL, 1, 1>(s, g0, &g1, use3m);
        launch_pair<T, 32, 32, GemmCfg<T>::BK_SMALL, 1, 1>(s, g2, nullptr, use3m);
        if (stop) RGF_HIP(hipEventRecord(stop, s));
    }
    RGF_HIP(hipGetLastError());
}

// gemm_args with the beta = 1 source X in place of C
template <typename T>
static GemmArgs<T> from_source(GemmArgs<T> g, CMat<T> X) {
    g.Xre = X.re;
    g.Xim = X.im;
    g.ldx = X.ld;
    return g;
}

template <typename T>
void cgemm(hipStream_t s, int64_t M, int64_t N, int64_t K, CMat<T> A, CMat<T> B, CMat<T> C, cd alpha,
           int beta, int64_t skip_lo, int64_t skip_hi, int use3m, T* ws, size_t ws_elems) {
    cgemm_skip<T>(s, M, N, K, A, B, C, alpha, beta, Skip{skip_lo, skip_hi, skip_lo, skip_hi}, use3m, ws, ws_elems);
}

// Split-K reduction of one forward-sweep product: D = sum_z P_z and, with
// X given, T = (float)(sc * X) + D (the next right-hand side; the same
// roundings as a conversion followed by an FP32 addition).
__global__ void sweep_reduce_kernel(int64_t M, int64_t N, int nz, const float* P, int64_t pstride, float* Dre,
                                    float* Dim, int64_t ldd, const double* Xre, const double* Xim, int64_t ldx,
                                    double sc, float* Tre, float* Tim, int64_t ldt) {
    const int64_t total = M * N;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        double sr = 0, si = 0;
        for (int z = 0; z < nz; ++z) {
            sr += (double)P[z * pstride + e];
            si += (double)P[z * pstride + total + e];
        }
        const int64_t i = e / N, j = e % N;
        const float dr = (float)sr, di = (float)si;
        Dre[i * ldd + j] = dr;
        Dim[i * ldd + j] = di;
        if (Xre) {
            Tre[i * ldt + j] = (float)(Xre[i * ldx + j] * sc) + dr;
            Tim[i * ldt + j] = (float)(Xim[i * ldx + j] * sc) + di;
        }
    }
}

void sweep_forward(hipStream_t s, int64_t N, int64_t m, CMat<float> G, CMat<float> t, CMat<float> D,
                   const CMat<double>* next, double sc, int use3m, float* ws, size_t ws_elems) {
    if (N <= 0 || m <= 0) return;
    // split-K as cgemm_skip would choose it, with the reduction fused
    const int64_t tiles = ((N + 63) / 64) * ((m + 63) / 64);
    const int64_t per = 2 * N * m;
    int64_t nz = 0, kc = 0;
    if (ws && N >= 256 && tiles < 512) nz = splitk_chunks<float>(tiles, N, ws_elems, per, kc);
    if (nz >= 2) {
        GemmArgs<float> g = gemm_args<float>(N, m, N, G, t, CMat<float>(ws, ws + N * m, m), cd(1, 0), 0, Skip{});
        g.nz = (int)nz;
        g.kc = (int)kc;
        g.zstride = per;
        if (launch_lean<float>(s, g, use3m) || launch_cplx<float, 64, 64, 16, 2, 2>(use3m, s, g)) {
            hipLaunchKernelGGL(sweep_reduce_kernel, dim3(grid_for(N * m)), dim3(256), 0, s, N, m, (int)nz,
                               (const float*)ws, per, D.re, D.im, D.ld, next ? next->re : nullptr,
                               next ? next->im : nullptr, next ? next->ld : 0, sc, t.re, t.im, t.ld);
            RGF_HIP(hipGetLastError());
            return;
        }
    }
    cgemm<float>(s, N, m, N, G, t, D, cd(1, 0), 0, 0, 0, use3m, ws, ws_elems);
    if (next) {
        convert<float, double>(s, N, m, *next, t, sc);
        add_mat<float>(s, N, m, t, D, t);
    }
}

// ------------------------------------------------------------------ K2: diagonal-block inverse
// One work group inverts the nb x nb diagonal block D = A[p0:p0+nb, p0:p0+nb]
// (nb <= NB, padded with the identity to NB) by Gauss-Jordan elimination with
// row interchanges confined to the block, writes D^-1 back to A and to Dout
// (ld = ldd), and folds the column interchanges that undo the row pivoting
// into the write-back.
// The nb x nb block (padded with the identity to NB) lives in LDS.  For
// each 16-column panel, all wavefronts run 16 Gauss-Jordan steps on the
// panel (each thread holds a few panel columns of one row in registers; the
// pivot row and column go through LDS, one barrier per step).  The other
// columns are then brought up to date: the 16 row interchanges, if any,
// then the rank-16 block sweep
//   M[:, J] = (rows P ? 0 : M[:, J]) + M'[:, P] * M_old[P, J]
// on MFMA, where M'[:, P] are the panel's new values.
template <typename T>
__device__ __forceinline__ T lane_bcast(T v, int l);
template <>
__device__ __forceinline__ float lane_bcast<float>(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
template <>
__device__ __forceinline__ double lane_bcast<double>(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// 64-bit maximum over a wavefront (DPP-free butterfly through ds_swizzle is
// not needed: the key reductions happen once per pivot or once per panel).
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const unsigned long long o = __shfl_xor(v, off);
        if (o > v) v = o;
    }
    return v;
}

// Value of lane 16 * (lane / 16) + k (the same lane group, column k): the
// DPP row broadcast (row_newbcast:k, a VALU move) instead of a permute
// through the LDS unit, whose latency sits on the serial chain of the
// pivot-tile elimination.
template <int K>
__device__ __forceinline__ int bcast16_bits(int v) {
    return __builtin_amdgcn_update_dpp(0, v, 0x150 + K, 0xf, 0xf, false);
}
template <int K>
__device__ __forceinline__ float bcast16(float v) {
    return __int_as_float(bcast16_bits<K>(__float_as_int(v)));
}
template <int K>
__device__ __forceinline__ double bcast16(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = bcast16_bits<K>((int)(b & 0xffffffffLL));
    const int hi = bcast16_bits<K>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// 1 / (ar + i ai): FP32 through the hardware reciprocal (the FP32 inverses
// only precondition the FP64 refinement), FP64 by IEEE division.
__device__ __forceinline__ void cinv(float ar, float ai, float& ivr, float& ivi) {
    const float r = __builtin_amdgcn_rcpf(ar * ar + ai * ai);
    ivr = ar * r;
    ivi = -ai * r;
}
__device__ __forceinline__ void cinv(double ar, double ai, double& ivr, double& ivi) {
    const double den = ar * ar + ai * ai;
    ivr = ar / den;
    ivi = -ai / den;
}

// One Gauss-Jordan step (pivot k) of the in-wavefront inverse of a 16x16
// complex tile held in the MFMA accumulator layout (lane (g, c) = (l/16, l%16)
// holds column c, rows Mfma<T>::row(g, r)), without row interchanges.  The
// pivot row comes by ds_bpermute from lane group grp(k), the multipliers
// M[i][k] by a DPP row broadcast from lane c = k of the own group, the pivot by
// readlane.  bad is set on a zero or non-finite pivot; gm tracks the largest
// |x|^2 of the new values (the growth of the unpivoted elimination).
template <typename T, int K>
__device__ __forceinline__ void gj_step16(typename Mfma<T>::v4& xr, typename Mfma<T>::v4& xi, int g, int c,
                                          bool& bad, float& gm) {
    constexpr int gk = Mfma<T>::grp(K), rk = Mfma<T>::elt(K);
    const T prr = __shfl(xr[rk], 16 * gk + c);
    const T pri = __shfl(xi[rk], 16 * gk + c);
    T fr[4], fi[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        fr[r] = bcast16<K>(xr[r]);
        fi[r] = bcast16<K>(xi[r]);
    }
    const T ar = lane_bcast<T>(xr[rk], 16 * gk + K);
    const T ai = lane_bcast<T>(xi[rk], 16 * gk + K);
    const T den = ar * ar + ai * ai;
    if (!(den > T(0)) || !isfinite(den)) bad = true;
    T ivr, ivi;
    cinv(ar, ai, ivr, ivi);
    const bool ck = c == K;
    const T scr = ck ? ivr : prr * ivr - pri * ivi;  // pivot row / pivot
    const T sci = ck ? ivi : prr * ivi + pri * ivr;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const T gr = fr[r] * ivr - fi[r] * ivi, gi = fr[r] * ivi + fi[r] * ivr;  // M[i][k] / pivot
        T nr = ck ? -gr : xr[r] - (fr[r] * scr - fi[r] * sci);
        T ni = ck ? -gi : xi[r] - (fr[r] * sci + fi[r] * scr);
        if (r == rk) {
            const bool pr = g == gk;  // this element is in the pivot row
            nr = pr ? scr : nr;
            ni = pr ? sci : ni;
        }
        xr[r] = nr;
        xi[r] = ni;
        gm = fmaxf(gm, (float)(nr * nr + ni * ni));
    }
}

// MFMA step s of acc += sgn * A * B for 16x16 complex tiles: A row-major in
// LDS (row stride ld, planes Ar/Ai), B in registers in the accumulator
// layout.  The K index of step s at lane group g is Mfma<T>::row(g, s), which
// makes the B tile's own element s the right operand.
template <typename T>
__device__ __forceinline__ void mma_step16(typename Mfma<T>::v4& accr, typename Mfma<T>::v4& acci, const T* Ar,
                                           const T* Ai, int ld, T br, T bi, T sgn, int g, int c, int s) {
    const int k = Mfma<T>::row(g, s);
    const T ar = sgn * Ar[c * ld + k], ai = sgn * Ai[c * ld + k];
    accr = Mfma<T>::mma(ar, br, accr);
    acci = Mfma<T>::mma(ar, bi, acci);
    accr = Mfma<T>::mma(-ai, bi, accr);
    acci = Mfma<T>::mma(ai, br, acci);
}
template <typename T>
__device__ __forceinline__ void mma_tile16(typename Mfma<T>::v4& accr, typename Mfma<T>::v4& acci, const T* Ar,
                                           const T* Ai, int ld, const typename Mfma<T>::v4& br,
                                           const typename Mfma<T>::v4& bi, T sgn, int g, int c) {
#pragma unroll
    for (int s = 0; s < 4; ++s) mma_step16<T>(accr, acci, Ar, Ai, ld, br[s], bi[s], sgn, g, c, s);
}

// acc -= A * B for 16x16 complex tiles, the product summed from zero and
// added once (one rounding at the scale of acc instead of one per MFMA step)
template <typename T>
__device__ __forceinline__ void mma_tile16_sub(typename Mfma<T>::v4& accr, typename Mfma<T>::v4& acci, const T* Ar,
                                               const T* Ai, int ld, const typename Mfma<T>::v4& br,
                                               const typename Mfma<T>::v4& bi, int g, int c) {
    typename Mfma<T>::v4 pr = {0, 0, 0, 0}, pi = {0, 0, 0, 0};
    mma_tile16<T>(pr, pi, Ar, Ai, ld, br, bi, T(1), g, c);
    accr -= pr;
    acci -= pi;
}

// Compile-time loop helper: f(std::integral_constant<int, K>) for K in [0, N).
template <int K, int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (K < N) {
        f(std::integral_constant<int, K>{});
        static_for<K + 1, N>(f);
    }
}

// The whole work group (NB / 16 wavefronts) takes part.
template <typename T, int NB>
__device__ __forceinline__ void gj_diag_blk(T* Are, T* Aim, int64_t lda, int p0, int nb, T* Dre, T* Dim,
                                            int64_t ldd, int* status) {
    using V = typename Mfma<T>::v4;
    constexpr int NW = NB / 16;  // wavefronts; wavefront w owns block column w
    constexpr int NT = NW * 64;
    constexpr int PW = 16;       // panel width
    constexpr int LDP = NB + 1;
    constexpr int NBLK = NB / 16;
    static_assert(NB % 64 == 0, "NB must be a multiple of 64");
    __shared__ T sr[NB][LDP], si[NB][LDP];
    __shared__ T orr[PW][NB], ori[PW][NB];  // M_old[P, :] for the block sweep
    constexpr int TPRW = NT / NB;    // phase-1 threads per row
    constexpr int CP = PW / TPRW;    // phase-1 panel columns per thread
    static_assert(TPRW * CP == PW, "bad phase-1 tiling");
    __shared__ int pivs[NB];
    __shared__ int bad_s, any_swap;
    __shared__ T colr[2][NB], coli[2][NB], rowr[2][PW], rowi[2][PW], pvr[PW], pvi[PW];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;

    // ---- fast path: no row interchanges, the matrix in registers.  Block
    // column w belongs to wavefront w.  Per panel p (Q = p + 1): after the
    // column panel A[:, P] and D_P = A_PP^-1 are published in LDS, every
    // wavefront w != p forms R_w = D_P A[P, w]; the owner of Q also brings its
    // pivot tile up to date and puts its column in LDS as the next column
    // panel (old tiles, R_Q in the rows P, the updated pivot tile).  After a
    // barrier the owner of Q inverts its pivot tile (D_Q) while every other
    // wavefront w updates its own tiles, A[I, w] -= A[I, P] R_w, and the tile
    // (w, Q) of the next panel in LDS, A[w, Q] -= A[w, P] R_Q: the owner's
    // elimination runs beside all the trailing tile updates.  The owner of P
    // replaces its column by -A[I, P] D_P and D_P.  If a panel fails the
    // growth checks (max|D| max|A[:, P]|, and the largest intermediate value
    // of the tile elimination over max|A_PP|, above 31.6 in FP32 or 100 in
    // FP64) or meets a zero pivot, the block is factored again below with
    // partial pivoting.
    {
        const int g = lane >> 4, c = lane & 15;
        const int jc = wave * 16 + c;
        // the block (identity padding outside nb) through LDS: coalesced rows
        // (thread t: column t % NB of rows t / NB + 4k), eight loads in flight
        // per thread before their LDS stores, then each wavefront takes its
        // block column
        {
            static_assert(NT % NB == 0 && (NB * NB / NT) % 8 == 0, "bad block load tiling");
            constexpr int RS = NT / NB, PER = NB * NB / NT;
            const int j = tid % NB, i0 = tid / NB;
#pragma unroll
            for (int k0 = 0; k0 < PER; k0 += 8) {
                T vr[8], vi[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int i = i0 + (k0 + u) * RS;
                    vr[u] = (i == j) ? T(1) : T(0);
                    vi[u] = 0;
                    if (i < nb && j < nb) {
                        vr[u] = Are[(int64_t)(p0 + i) * lda + p0 + j];
                        vi[u] = Aim[(int64_t)(p0 + i) * lda + p0 + j];
                    }
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    sr[i0 + (k0 + u) * RS][j] = vr[u];
                    si[i0 + (k0 + u) * RS][j] = vi[u];
                }
            }
        }
        __syncthreads();
        V xr[NW], xi[NW];
#pragma unroll
        for (int bi = 0; bi < NW; ++bi)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                xr[bi][r] = sr[bi * 16 + Mfma<T>::row(g, r)][jc];
                xi[bi][r] = si[bi * 16 + Mfma<T>::row(g, r)][jc];
            }
        __syncthreads();
        // LDS views (aliasing sr / si): column panel [2][NB][CLD], D [2][16][CLD]
        // (rows of CLD elements, 16-byte multiples, so the four K values an
        // MFMA lane takes are one aligned read)
        constexpr int CLD = 20;
        static_assert(2 * NB * CLD + 2 * 16 * CLD <= NB * LDP, "fast-path buffers exceed LDS");
        T* const fr = &sr[0][0];
        T* const fi = &si[0][0];
        auto cbuf = [&](int buf, int plane) { return (plane ? fi : fr) + buf * NB * CLD; };
        auto dbuf = [&](int buf, int plane) { return (plane ? fi : fr) + 2 * NB * CLD + buf * 16 * CLD; };
        // per panel: elimination verdict, max|D|^2, and max|x|^2 of each tile of
        // its column panel (written by the wavefront that forms the tile); the
        // growth checks are evaluated once, after the last panel
        __shared__ int okp[NBLK];
        __shared__ float dmax[NBLK], tmax[NBLK][NBLK];
        auto tile_max = [&](const V& vr, const V& vi) {
            float m = 0.f;
#pragma unroll
            for (int r = 0; r < 4; ++r) m = fmaxf(m, (float)(vr[r] * vr[r] + vi[r] * vi[r]));
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
            return m;
        };
        // tile (bi, column) <-> column panel buffer in the accumulator layout
        auto put_tile = [&](int buf, int bi, const V& vr, const V& vi) {
            T* const cr = cbuf(buf, 0) + bi * 16 * CLD;
            T* const ci = cbuf(buf, 1) + bi * 16 * CLD;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                cr[Mfma<T>::row(g, r) * CLD + c] = vr[r];
                ci[Mfma<T>::row(g, r) * CLD + c] = vi[r];
            }
        };
        auto get_tile = [&](int buf, int bi, V& vr, V& vi) {
            const T* const cr = cbuf(buf, 0) + bi * 16 * CLD;
            const T* const ci = cbuf(buf, 1) + bi * 16 * CLD;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                vr[r] = cr[Mfma<T>::row(g, r) * CLD + c];
                vi[r] = ci[Mfma<T>::row(g, r) * CLD + c];
            }
        };
        const float lim = sizeof(T) == 4 ? 1e3f : 1e4f;  // squared growth bounds
        // D of the panel this wavefront owns (kept from its elimination)
        V dr = V{0, 0, 0, 0}, di = V{0, 0, 0, 0};
        // Invert the pivot tile of panel q (in xr[q], up to date) into dr/di;
        // publish D_q, max|D_q|^2 and the verdict of the elimination checks.
        auto factor = [&](int q) {
            // the elimination is the panel's critical path: its wave takes
            // issue priority over the tile-update wave sharing its SIMD
            __builtin_amdgcn_s_setprio(2);
#pragma unroll
            for (int bi = 0; bi < NW; ++bi)
                if (bi == q) {
                    dr = xr[bi];
                    di = xi[bi];
                }
            float mt = 0.f;
#pragma unroll
            for (int r = 0; r < 4; ++r) mt = fmaxf(mt, (float)(dr[r] * dr[r] + di[r] * di[r]));
            bool bad = false;
            float mg = 0.f;
            static_for<0, 16>([&](auto kc) { gj_step16<T, decltype(kc)::value>(dr, di, g, c, bad, mg); });
            const int buf = q & 1;
            float md = 0.f;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                dbuf(buf, 0)[Mfma<T>::row(g, r) * CLD + c] = dr[r];
                dbuf(buf, 1)[Mfma<T>::row(g, r) * CLD + c] = di[r];
                md = fmaxf(md, (float)(dr[r] * dr[r] + di[r] * di[r]));
            }
            const float wd = __uint_as_float((unsigned)wave_max_u64(__float_as_uint(md)));
            const float wt = __uint_as_float((unsigned)wave_max_u64(__float_as_uint(mt)));
            const float wg = __uint_as_float((unsigned)wave_max_u64(__float_as_uint(mg)));
            // growth of the in-tile elimination max|x| / max|A_QQ|, squared
            const bool grow_ok = isfinite(wg) && wg <= lim * fmaxf(wt, 1e-30f);
            if (lane == 0) {
                okp[q] = (!bad && grow_ok) ? 1 : 0;
                dmax[q] = wd;
            }
            __builtin_amdgcn_s_setprio(0);
        };
        if (wave == 0) {
            float m = 0.f;
#pragma unroll
            for (int bi = 0; bi < NW; ++bi) {
                put_tile(0, bi, xr[bi], xi[bi]);
                m = fmaxf(m, tile_max(xr[bi], xi[bi]));
            }
            if (lane == 0)
                for (int bi = 0; bi < NW; ++bi) tmax[0][bi] = m;
            factor(0);
        }
        // A panel that fails the checks only makes the later panels produce
        // values that are discarded: nothing is written before the verdict.
        for (int p = 0; p < NBLK; ++p) {
            const int buf = p & 1, q = p + 1, nbuf = q & 1;
            const T* const cr = cbuf(buf, 0);
            const T* const ci = cbuf(buf, 1);
            __syncthreads();  // panel p's column and D are published
            V rr = V{0, 0, 0, 0}, ri = V{0, 0, 0, 0};
            if (wave == p) {
                // own column: -A[I, P] D, and D in the pivot rows
#pragma unroll
                for (int bi = 0; bi < NW; ++bi) {
                    if (bi == p) {
                        xr[bi] = dr;
                        xi[bi] = di;
                    } else {
                        V nr = V{0, 0, 0, 0}, ni = V{0, 0, 0, 0};
                        mma_tile16<T>(nr, ni, cr + bi * 16 * CLD, ci + bi * 16 * CLD, CLD, dr, di, T(-1), g, c);
                        xr[bi] = nr;
                        xi[bi] = ni;
                    }
                }
            } else {
                // R = D A[P, w], stored in place of A[P, w]
                V br = xr[0], bim = xi[0];
#pragma unroll
                for (int bi = 1; bi < NW; ++bi)
                    if (bi == p) {
                        br = xr[bi];
                        bim = xi[bi];
                    }
                mma_tile16<T>(rr, ri, dbuf(buf, 0), dbuf(buf, 1), CLD, br, bim, T(1), g, c);
#pragma unroll
                for (int bi = 0; bi < NW; ++bi)
                    if (bi == p) {
                        xr[bi] = rr;
                        xi[bi] = ri;
                    }
                if (wave == q) {
                    // the next column panel: old tiles, R_Q in the rows P, and
                    // the pivot tile brought up to date here
#pragma unroll
                    for (int bi = 0; bi < NW; ++bi) {
                        if (bi == q)
                            mma_tile16_sub<T>(xr[bi], xi[bi], cr + bi * 16 * CLD, ci + bi * 16 * CLD, CLD, rr, ri, g, c);
                        put_tile(nbuf, bi, xr[bi], xi[bi]);
                    }
                    float mp = 0.f, mq = 0.f;
#pragma unroll
                    for (int bi = 0; bi < NW; ++bi) {
                        if (bi == p) mp = tile_max(xr[bi], xi[bi]);
                        if (bi == q) mq = tile_max(xr[bi], xi[bi]);
                    }
                    if (lane == 0) {
                        tmax[q][p] = mp;
                        tmax[q][q] = mq;
                    }
                }
            }
            if (q == NBLK) {
                // last panel: the remaining tile updates, no next panel
                if (wave != p) {
#pragma unroll
                    for (int bi = 0; bi < NW; ++bi)
                        if (bi != p)
                            mma_tile16_sub<T>(xr[bi], xi[bi], cr + bi * 16 * CLD, ci + bi * 16 * CLD, CLD, rr, ri, g, c);
                }
                break;
            }
            __syncthreads();  // the next column panel's old tiles and R_Q are in LDS
            if (wave == q) {
                factor(q);
            } else if (wave != p) {
#pragma unroll
                for (int bi = 0; bi < NW; ++bi)
                    if (bi != p)
                        mma_tile16_sub<T>(xr[bi], xi[bi], cr + bi * 16 * CLD, ci + bi * 16 * CLD, CLD, rr, ri, g, c);
                // tile (w, Q) of the next panel: A[w, Q] -= A[w, P] R_Q
                V tr, ti, qr, qi;
                get_tile(nbuf, wave, tr, ti);
                get_tile(nbuf, p, qr, qi);
                mma_tile16_sub<T>(tr, ti, cr + wave * 16 * CLD, ci + wave * 16 * CLD, CLD, qr, qi, g, c);
                put_tile(nbuf, wave, tr, ti);
                const float m = tile_max(tr, ti);
                if (lane == 0) tmax[q][wave] = m;
            }
        }
        __syncthreads();  // every verdict is in LDS; the panel buffers (aliasing sr / si) are free
        // checks, squared: multipliers max|A[:, P]| max|D_P| and the growth of
        // the in-tile elimination stay below 31.6 (FP32) or 100 (FP64)
        bool ok = true;
        for (int p = 0; p < NBLK; ++p) {
            float m = 0.f;
            for (int bi = 0; bi < NBLK; ++bi) m = fmaxf(m, tmax[p][bi]);
            const float gr = m * dmax[p];
            ok = ok && okp[p] && isfinite(gr) && gr <= lim;
        }
        if (ok) {
            // D^-1 from the registers to A and to Dout
#pragma unroll
            for (int bi = 0; bi < NW; ++bi)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int i = bi * 16 + Mfma<T>::row(g, r);
                    if (i < nb && jc < nb) {
                        Are[(int64_t)(p0 + i) * lda + p0 + jc] = xr[bi][r];
                        Aim[(int64_t)(p0 + i) * lda + p0 + jc] = xi[bi][r];
                        Dre[(int64_t)i * ldd + jc] = xr[bi][r];
                        Dim[(int64_t)i * ldd + jc] = xi[bi][r];
                    }
                }
            return;
        }
    }
    __syncthreads();

    // ---- fallback: the block again from A, in LDS, with partial pivoting
    const int prow = tid / TPRW, pc0 = (tid % TPRW) * CP;
    if (tid == 0) bad_s = 0;
    for (int e = tid; e < NB * NB; e += NT) {
        const int i = e / NB, j = e % NB;
        T vr = (i == j) ? T(1) : T(0), vi = 0;
        if (i < nb && j < nb) {
            vr = Are[(int64_t)(p0 + i) * lda + p0 + j];
            vi = Aim[(int64_t)(p0 + i) * lda + p0 + j];
        }
        sr[i][j] = vr;
        si[i][j] = vi;
    }
    __syncthreads();
    for (int q0 = 0; q0 < NB; q0 += PW) {
        // ---- phase 1: all threads factor the panel [q0, q0+16) with partial
        // pivoting over the rows >= the pivot row: thread t owns CP panel
        // columns of row t / TPRW.  Each pivot step publishes the pivot row
        // and the pivot column through LDS (double-buffered, one barrier per
        // step, one more when rows are interchanged).
        {
            T xr[CP], xi[CP];
#pragma unroll
            for (int c = 0; c < CP; ++c) {
                xr[c] = sr[prow][q0 + pc0 + c];
                xi[c] = si[prow][q0 + pc0 + c];
            }
            int bad = 0;
            bool swapped = false;
#pragma unroll
            for (int k = 0; k < PW; ++k) {
                const int kc = q0 + k, bsel = k & 1;
                // publish column k (its owner in every row) and row kc
                if ((tid % TPRW) == k / CP) {
                    colr[bsel][prow] = xr[k % CP];
                    coli[bsel][prow] = xi[k % CP];
                }
                if (prow == kc) {
#pragma unroll
                    for (int c = 0; c < CP; ++c) {
                        rowr[bsel][pc0 + c] = xr[c];
                        rowi[bsel][pc0 + c] = xi[c];
                    }
                }
                __syncthreads();
                // every wavefront finds the same pivot: largest |x| among rows >= kc
                unsigned long long key = 0;
                for (int rr2 = kc + lane; rr2 < NB; rr2 += 64) {
                    const T a = colr[bsel][rr2], b = coli[bsel][rr2];
                    const float m = (float)(a * a + b * b);
                    const unsigned long long kk = ((unsigned long long)__float_as_uint(m) << 32) |
                                                  (unsigned long long)(0xffffffffu - (unsigned)rr2);
                    if (kk > key) key = kk;
                }
                key = wave_max_u64(key);
                const int piv = (int)(0xffffffffu - (unsigned)(key & 0xffffffffu));
                const bool xch = piv != kc;  // uniform
                if (xch) {
                    swapped = true;
                    if (prow == piv) {
#pragma unroll
                        for (int c = 0; c < CP; ++c) {
                            pvr[pc0 + c] = xr[c];
                            pvi[pc0 + c] = xi[c];
                        }
                    }
                    __syncthreads();
                }
                const T pkr = xch ? pvr[k] : rowr[bsel][k];
                const T pki = xch ? pvi[k] : rowi[bsel][k];
                const T den = pkr * pkr + pki * pki;
                if (!(den > 0) || !isfinite(den)) bad = 1;
                const T ivr = pkr / den, ivi = -pki / den;
                if (prow == kc) {
#pragma unroll
                    for (int c = 0; c < CP; ++c) {
                        const int j = pc0 + c;
                        const T br = xch ? pvr[j] : rowr[bsel][j], bi2 = xch ? pvi[j] : rowi[bsel][j];
                        xr[c] = (j == k) ? ivr : br * ivr - bi2 * ivi;
                        xi[c] = (j == k) ? ivi : br * ivi + bi2 * ivr;
                    }
                } else {
                    // row piv receives the old row kc before its elimination
                    const bool moved = xch && prow == piv;
                    const T fr = moved ? colr[bsel][kc] : colr[bsel][prow];
                    const T fi = moved ? coli[bsel][kc] : coli[bsel][prow];
                    const T gr = fr * ivr - fi * ivi, gi = fr * ivi + fi * ivr;  // f / pivot
#pragma unroll
                    for (int c = 0; c < CP; ++c) {
                        const int j = pc0 + c;
                        T br = moved ? rowr[bsel][j] : xr[c];
                        T bi2 = moved ? rowi[bsel][j] : xi[c];
                        const T yr = xch ? pvr[j] : rowr[bsel][j], yi = xch ? pvi[j] : rowi[bsel][j];
                        if (j == k) {
                            xr[c] = -gr;
                            xi[c] = -gi;
                        } else {
                            xr[c] = br - (gr * yr - gi * yi);
                            xi[c] = bi2 - (gr * yi + gi * yr);
                        }
                    }
                }
                if (tid == 0) pivs[kc] = piv;
            }
            if (bad) bad_s = 1;
#pragma unroll
            for (int c = 0; c < CP; ++c) {
                sr[prow][q0 + pc0 + c] = xr[c];
                si[prow][q0 + pc0 + c] = xi[c];
            }
            if (tid == 0) any_swap = swapped ? 1 : 0;
        }
        __syncthreads();
        // ---- phase 2a: row interchanges on the other columns; keep M_old[P, :]
        for (int j = tid; j < NB; j += NT) {
            if (j >= q0 && j < q0 + PW) continue;
            for (int k = 0; any_swap && k < PW; ++k) {
                const int a = q0 + k, b = pivs[a];
                if (b != a) {
                    const T tr = sr[a][j], ti = si[a][j];
                    sr[a][j] = sr[b][j];
                    si[a][j] = si[b][j];
                    sr[b][j] = tr;
                    si[b][j] = ti;
                }
            }
#pragma unroll
            for (int k = 0; k < PW; ++k) {
                orr[k][j] = sr[q0 + k][j];
                ori[k][j] = si[q0 + k][j];
            }
        }
        __syncthreads();
        // ---- phase 2b: rank-16 block sweep of the other columns on MFMA
        for (int t = wave; t < NBLK * NBLK; t += NT / 64) {
            const int bi = t / NBLK, bj = t % NBLK;
            if (bj * 16 == q0) continue;
            const int l16 = lane & 15, l4 = lane >> 4;
            V a0, a1;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = bi * 16 + Mfma<T>::row(l4, r), j = bj * 16 + l16;
                const bool inP = (i >= q0 && i < q0 + PW);
                a0[r] = inP ? T(0) : sr[i][j];
                a1[r] = inP ? T(0) : si[i][j];
            }
#pragma unroll
            for (int kk = 0; kk < PW; kk += 4) {
                const T ar = sr[bi * 16 + l16][q0 + kk + l4], ai = si[bi * 16 + l16][q0 + kk + l4];
                const T br = orr[kk + l4][bj * 16 + l16], bi2 = ori[kk + l4][bj * 16 + l16];
                a0 = Mfma<T>::mma(ar, br, a0);
                a1 = Mfma<T>::mma(ar, bi2, a1);
                a0 = Mfma<T>::mma(-ai, bi2, a0);
                a1 = Mfma<T>::mma(ai, br, a1);
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = bi * 16 + Mfma<T>::row(l4, r), j = bj * 16 + l16;
                sr[i][j] = a0[r];
                si[i][j] = a1[r];
            }
        }
        __syncthreads();
    }
    // column un-permutation: out[:, j] = X[:, src[j]]; pivs becomes the
    // destination column of each X column
    __shared__ int srcs[NB];
    if (tid == 0) {
        for (int j = 0; j < NB; ++j) srcs[j] = j;
        for (int k = NB - 1; k >= 0; --k) {
            const int p = pivs[k];
            const int t2 = srcs[k];
            srcs[k] = srcs[p];
            srcs[p] = t2;
        }
        for (int j = 0; j < NB; ++j) pivs[srcs[j]] = j;
    }
    __syncthreads();
    for (int e = tid; e < NB * NB; e += NT) {
        const int i = e / NB, c = e % NB;
        const int j = pivs[c];
        if (i < nb && j < nb) {
            Are[(int64_t)(p0 + i) * lda + p0 + j] = sr[i][c];
            Aim[(int64_t)(p0 + i) * lda + p0 + j] = si[i][c];
            Dre[(int64_t)i * ldd + j] = sr[i][c];
            Dim[(int64_t)i * ldd + j] = si[i][c];
        }
    }
    if (tid == 0 && bad_s) atomicOr(status, 1);
}

template <typename T, int NB>
__global__ __launch_bounds__(NB / 16 * 64) void gj_diag_blk_kernel(T* Are, T* Aim, int64_t lda, int p0, int nb,
                                                                   T* Dre, T* Dim, int64_t ldd, int* status) {
    gj_diag_blk<T, NB>(Are, Aim, lda, p0, nb, Dre, Dim, ldd, status);
}

// Partial write-back of block [p0, p0+nb): with rows != 0, the row panel
// A[P, j] = R[:, j] (j outside P); and the column panel A[i, P] = C[i, :] for
// rows i in [clo, chi) outside [slo, shi).
template <typename T>
__global__ void gj_writeback2_kernel(T* Are, T* Aim, int64_t lda, int64_t N, int p0, int nb, const T* Rre,
                                     const T* Rim, int64_t ldr, const T* Cre, const T* Cim, int64_t ldc, int rows,
                                     int64_t clo, int64_t chi, int64_t slo, int64_t shi) {
    const int64_t nrow = rows ? (int64_t)nb * N : 0;
    const int64_t ncol = (chi - clo) * nb;
    const int64_t total = nrow > ncol ? nrow : ncol;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        if (e < nrow) {
            const int64_t i = e / N, j = e % N;
            if (j < p0 || j >= p0 + nb) {
                Are[(p0 + i) * lda + j] = Rre[i * ldr + j];
                Aim[(p0 + i) * lda + j] = Rim[i * ldr + j];
            }
        }
        if (e < ncol) {
            const int64_t i = clo + e / nb, j = e % nb;
            if (i < slo || i >= shi) {
                Are[i * lda + p0 + j] = Cre[i * ldc + j];
                Aim[i * lda + p0 + j] = Cim[i * ldc + j];
            }
        }
    }
}

// Copies of the block column A[:, p0:p0+nb] (to C, ld ldc) and block row
// A[p0:p0+nb, :] (to R, ld ldr), both planes; *zero (if given) is cleared.
template <typename T>
__global__ void panel_copy_kernel(const T* Are, const T* Aim, int64_t lda, int64_t N, int64_t p0, int nb, T* Cre,
                                  T* Cim, int64_t ldc, T* Rre, T* Rim, int64_t ldr, unsigned* zero) {
    const int64_t half = (int64_t)nb * N;
    if (zero && blockIdx.x == 0 && threadIdx.x == 0) *zero = 0;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < 2 * half;
         e += (int64_t)gridDim.x * blockDim.x) {
        if (e < half) {
            const int64_t i = e / nb, j = e % nb;
            Cre[i * ldc + j] = Are[i * lda + p0 + j];
            Cim[i * ldc + j] = Aim[i * lda + p0 + j];
        } else {
            const int64_t f = e - half, i = f / N, j = f % N;
            Rre[i * ldr + j] = Are[(p0 + i) * lda + j];
            Rim[i * ldr + j] = Aim[(p0 + i) * lda + j];
        }
    }
}

template <typename T>
static constexpr int diag_bmax() {
    return diag_block_max<T>();  // the diagonal block lives in LDS
}

void KernelTimer::begin(hipStream_t s) {
    active = (seen++ % std::max(stride, 1)) == 0;
    if (!active) return;
    if (used + 2 > ev.size()) {
        for (int q = 0; q < 64; ++q) {
            hipEvent_t e;
            RGF_HIP(hipEventCreateWithFlags(&e, hipEventReleaseToDevice));
            ev.push_back(e);
        }
    }
    RGF_HIP(hipEventRecord(ev[used], s));
}

void KernelTimer::end(hipStream_t s, double kernel_flops) {
    if (!active) return;
    active = false;
    RGF_HIP(hipEventRecord(ev[used + 1], s));
    used += 2;
    flops += kernel_flops;
    launches += 1;
}

double KernelTimer::seconds() {
    double t = 0;
    for (size_t q = 0; q + 1 < used; q += 2) {
        RGF_HIP(hipEventSynchronize(ev[q + 1]));
        float ms = 0;
        RGF_HIP(hipEventElapsedTime(&ms, ev[q], ev[q + 1]));
        t += ms * 1e-3;
    }
    return t;
}

KernelTimer::~KernelTimer() {
    for (auto e : own) (void)hipEventDestroy(e);
}

template <typename T>
size_t inverse_workspace_elems(int64_t N, int b) {
    // two sets of R (b x N) and C (N x b), plus D (b x b); two planes each
    return (size_t)2 * (2 * (size_t)2 * b * N + (size_t)b * b);
}

template <typename T>
static void launch_diag(hipStream_t s, CMat<T> A, int64_t p0, int nb, CMat<T> D, int* st) {
    if constexpr (diag_bmax<T>() >= 128) {
        if (nb > 64) {
            hipLaunchKernelGGL((gj_diag_blk_kernel<T, 128>), dim3(1), dim3(512), 0, s, A.re, A.im, A.ld, (int)p0, nb,
                               D.re, D.im, D.ld, st);
            return;
        }
    }
    hipLaunchKernelGGL((gj_diag_blk_kernel<T, 64>), dim3(1), dim3(256), 0, s, A.re, A.im, A.ld, (int)p0, nb, D.re,
                       D.im, D.ld, st);
}

// Blocked Gauss-Jordan (sweep operator) in place.  For pivot block P:
//   D = A_PP^-1,  R = D A_P:,  C = -A_:P D,  A_JJ -= A_JP R_:J,  A_PJ = R_PJ,
//   A_JP = C_JP,  A_PP = D.
// With a second stream (look-ahead), the bulk update of step p runs there
// while the first stream forms the panels of block p+1 and factors block
// p+2, so the single-work-group diagonal inverse and the panel products
// overlap the bulk update.
template <typename T>
void invert_inplace(hipStream_t s, int64_t N, CMat<T> A, int b, T* ws, int* dev_status, int use3m,
                    KernelTimer* timer, const Overlap* ov) {
    if (N <= 0) return;
    if (b <= 0) b = diag_bmax<T>();
    if (b > diag_bmax<T>() || (b % 16) != 0) throw Error(RGF_ERR_INVALID_ARGUMENT, "invalid inverse block size");
    const size_t set = (size_t)2 * 2 * b * N;
    CMat<T> R[2], C[2];
    for (int q = 0; q < 2; ++q) {
        T* base = ws + q * set;
        R[q] = CMat<T>(base, base + (size_t)b * N, N);
        C[q] = CMat<T>(base + 2 * (size_t)b * N, base + 3 * (size_t)b * N, b);
    }
    CMat<T> D(ws + 2 * set, ws + 2 * set + (size_t)b * b, b);
    const int64_t nblk = (N + b - 1) / b;
    // R = D A_P: and C = -A_:P D in one launch (D from the diagonal block)
    // the largest multiplier |(A[:, P] D)_ij| of every step goes to the word
    // after the status flags (pivoting is confined to the diagonal block,
    // so this is the growth the rows outside the block see)
    unsigned* const amax = (unsigned*)(dev_status + 1);
    auto panel_products = [&](hipStream_t st, int64_t p0, int nb, int q) {
        const GemmArgs<T> gr = gemm_args<T>(nb, N, nb, D, A.offset(p0, 0), R[q], cd(1, 0), 0, Skip{});
        GemmArgs<T> gc = gemm_args<T>(N, nb, nb, A.offset(0, p0), D, C[q], cd(-1, 0), 0, Skip{});
        gc.amax = amax;
        cgemm_pair<T>(st, gr, gc, use3m);
    };
    auto panels = [&](hipStream_t st, int64_t p0, int nb, int q) {
        launch_diag<T>(st, A, p0, nb, D, dev_status);
        panel_products(st, p0, nb, q);
    };
    auto writeback = [&](hipStream_t st, int64_t p0, int nb, int q, int rows, int64_t clo, int64_t chi, int64_t slo,
                         int64_t shi) {
        const int64_t tot = std::max<int64_t>(rows ? (int64_t)nb * N : 0, (chi - clo) * nb);
        if (tot <= 0) return;
        hipLaunchKernelGGL((gj_writeback2_kernel<T>), dim3(grid_for(tot, 256, 4096)), dim3(256), 0, st, A.re, A.im,
                           A.ld, N, (int)p0, nb, R[q].re, R[q].im, R[q].ld, C[q].re, C[q].im, C[q].ld, rows, clo, chi,
                           slo, shi);
    };
    const double fl = use3m ? 6.0 : 8.0;
    if (nblk == 1) {
        launch_diag<T>(s, A, 0, (int)N, D, dev_status);
        RGF_HIP(hipGetLastError());
        return;
    }
    if (!ov || !ov->s2 || !ov->s3) {
        for (int64_t p0 = 0; p0 < N; p0 += b) {
            const int nb = (int)std::min<int64_t>(b, N - p0);
            panels(s, p0, nb, 0);
            if (timer) timer->begin(s);
            cgemm<T>(s, N, N, nb, A.offset(0, p0), R[0], A, cd(-1, 0), 1, p0, p0 + nb, use3m);
            if (timer) timer->end(s, fl * (double)(N - nb) * (double)(N - nb) * nb);
            writeback(s, p0, nb, 0, 1, 0, N, p0, p0 + nb);
        }
        RGF_HIP(hipGetLastError());
        return;
    }
    // Look-ahead on two streams.  Per block q (P = block q, Q = q+1, Q2 = q+2):
    //  second stream: one launch brings block row and column Q up to date
    //                 with step q (outside P and Q, written to the panel
    //                 copies Rold and Aold[Q]) and applies step q to the
    //                 block (Q2, Q2) in place; then the bulk update of step
    //                 q outside P, Q and (Q2, Q2);
    //  first stream : forms Q's panels in place from the copies (the blocks
    //                 A[Q, P] and A[P, Q] reached the copies with step q's
    //                 panels), applies step q+1 to (Q2, Q2) and queues its
    //                 inverse, which is dispatched beside the bulk update of
    //                 step q (its single work group needs a whole CU).
    // Every write to A is in place and no write-back follows the bulk update;
    // the second stream waits only for the panels of the step it applies.
    hipStream_t s2 = ov->s2, s3 = ov->s3;
    hipEvent_t evS[2] = {ov->ev[0], ov->ev[1]}, evT[2] = {ov->ev[2], ov->ev[3]};
    hipEvent_t evB[2] = {ov->ev[5], ov->ev[6]};
    CMat<T> Aold[2], Rold;
    {
        // workspace: Aold[0], Aold[1] (N x b), Rold (b x N), then D
        const size_t pl = (size_t)3 * b * N + (size_t)b * b;
        Aold[0] = CMat<T>(ws, ws + pl, b);
        Aold[1] = CMat<T>(ws + (size_t)b * N, ws + pl + (size_t)b * N, b);
        Rold = CMat<T>(ws + (size_t)2 * b * N, ws + pl + (size_t)2 * b * N, N);
        D = CMat<T>(ws + (size_t)3 * b * N, ws + pl + (size_t)3 * b * N, b);
    }
    // A 64-wide first block when b = 128 (all block edges stay on the 64-wide
    // tile grid): the first diagonal inverse, which nothing overlaps, costs
    // about a third of a 128-wide one.
    const int b0 = (b == 128 && N > 2 * b) ? 64 : b;
    const int64_t nbl = b0 < b ? 1 + (N - b0 + b - 1) / b : nblk;
    auto p_of = [&](int64_t q) { return q == 0 ? (int64_t)0 : b0 + (q - 1) * (int64_t)b; };
    auto nb_of = [&](int64_t q) {
        return q < nbl ? (int)std::min<int64_t>(q == 0 ? b0 : b, N - p_of(q)) : 0;
    };
    // A[P, J] = D Rold[:, J] and A[I, P] = -Aold[I, :] D outside P.  The
    // blocks the next copies need from these panels, A[Q, P] and A[P, Q], are
    // also stored to Rold[:, P] and Aold[Q][P, :] (Rold[:, P] is not read
    // here: the tiles of columns P are skipped).
    // A[K, K] -= Aold_j[K, :] A[Pj, K]: step j applied to diagonal block k;
    // `stop` is recorded when it completes
    auto schur = [&](int64_t k, int64_t j, hipEvent_t stop) {
        const int64_t pk = p_of(k), pj = p_of(j);
        const int nk = nb_of(k), nj = nb_of(j);
        const GemmArgs<T> g = gemm_args<T>(nk, nk, nj, Aold[j & 1].offset(pk, 0), A.offset(pj, pk), A.offset(pk, pk),
                                           cd(-1, 0), 1, Skip{});
        if (!launch_cplx<T, 32, 32, 16, 1, 1>(use3m, s, g, nullptr, nullptr, stop)) {
            cgemm_skip<T>(s, nk, nk, nj, Aold[j & 1].offset(pk, 0), A.offset(pj, pk), A.offset(pk, pk), cd(-1, 0), 1,
                          Skip{}, use3m, nullptr, 0);
            if (stop) RGF_HIP(hipEventRecord(stop, s));
        }
    };
    // FP32: the step on the block two ahead rides in the panel launch (one
    // launch and one launch gap fewer on the diagonal chain, which sets the
    // FP32 step time: 1.4-2.6 % off the inverse at N = 4096).  Its tiles come
    // last in the grid and wait for the row-panel tiles they read (the
    // columns of block q+1), which count themselves into a counter kept
    // after the panel copies; `handed` is the count the device reaches once
    // every launched hand-over tile is done.  FP64 keeps the separate launch
    // (its step time is the bulk update's; fused measured 0.3 % slower).
    unsigned* const hand = reinterpret_cast<unsigned*>(ws + 2 * ((size_t)3 * b * N + (size_t)b * b));
    unsigned handed = 0;
    auto panels_inplace = [&](int64_t q, hipEvent_t stop) {
        const int64_t p0 = p_of(q), p1 = p0 + nb_of(q);
        const int nb = nb_of(q), nb1 = nb_of(q + 1);
        GemmArgs<T> gr = gemm_args<T>(nb, N, nb, D, Rold, A.offset(p0, 0), cd(1, 0), 0, Skip{0, 0, p0, p1});
        GemmArgs<T> gc = gemm_args<T>(N, nb, nb, Aold[q & 1], D, A.offset(0, p0), cd(-1, 0), 0, Skip{p0, p1, 0, 0});
        gc.amax = amax;
        if (nb1 > 0) {
            const CMat<T> An = Aold[(q + 1) & 1].offset(p0, 0), Rp = Rold.offset(0, p0);
            gr.Yre = An.re;
            gr.Yim = An.im;
            gr.ldy = An.ld;
            gr.y_rlo = 0;
            gr.y_rhi = nb;
            gr.y_clo = (int)p1;
            gr.y_chi = (int)(p1 + nb1);
            gc.Yre = Rp.re;
            gc.Yim = Rp.im;
            gc.ldy = Rp.ld;
            gc.y_rlo = (int)p1;
            gc.y_rhi = (int)(p1 + nb1);
            gc.y_clo = 0;
            gc.y_chi = nb;
        }
        if (nb1 > 0 && sizeof(T) == 8) {
            cgemm_pair<T>(s, gr, gc, use3m);
            schur(q + 1, q, stop);
            return;
        }
        if (nb1 > 0) {
            // step q on block q+1: A[Q, Q] -= Aold_q[Q, :] A[P, Q], A[P, Q] from this launch's row panel
            constexpr int BK = GemmCfg<T>::BK2;
            GemmArgs<T> gs = gemm_args<T>(nb1, nb1, nb, Aold[q & 1].offset(p1, 0), A.offset(p0, p1), A.offset(p1, p1),
                                          cd(-1, 0), 1, Skip{});
            unsigned count = 0;  // row-panel tiles (64 x 64) meeting the columns of block q+1
            for (int64_t t = 0; t < N; t += 64)
                if (t < p1 + nb1 && t + 64 > p1 && !(t >= p0 && t + 64 <= p1)) count += (unsigned)((nb + 63) / 64);
            GemmArgs<T> gh = gr;
            gh.ctr = hand;
            gh.ctr_clo = (int)p1;
            gh.ctr_chi = (int)(p1 + nb1);
            gs.ctr = hand;
            gs.ctr_wait = handed + count;
            gs.fail = dev_status;
            if (launch_cplx<T, 64, 64, BK, 2, 2>(use3m, s, gh, &gc, &gs, stop)) {
                handed += count;
            } else {
                cgemm_pair<T>(s, gr, gc, use3m);
                schur(q + 1, q, stop);
            }
            RGF_HIP(hipGetLastError());
            return;
        }
        cgemm_pair<T>(s, gr, gc, use3m, stop);
    };
    {
        // step 0 panels from full copies of block row and column 0
        const int nb = nb_of(0);
        const int64_t tot = 2 * (int64_t)nb * N;
        launch_diag<T>(s, A, 0, nb, D, dev_status);
        hipLaunchKernelGGL((panel_copy_kernel<T>), dim3(grid_for(tot, 256, 4096)), dim3(256), 0, s, A.re, A.im, A.ld,
                           N, (int64_t)0, nb, Aold[0].re, Aold[0].im, Aold[0].ld, Rold.re, Rold.im, Rold.ld, hand);
        panels_inplace(0, evS[0]);  // with step 0 on block 1
        launch_diag<T>(s, A, p_of(1), nb_of(1), D, dev_status);
    }
    for (int64_t q = 0; q < nbl; ++q) {
        const int64_t p0 = p_of(q), p1 = p_of(q + 1);
        const int nb = nb_of(q), nb1 = nb_of(q + 1), nb2 = nb_of(q + 2);
        const int64_t p2 = p1 + nb1, p3 = p2 + nb2;
        CMat<T> An = Aold[(q + 1) & 1];
        RGF_HIP(hipStreamWaitEvent(s2, evS[q & 1], 0));
        if (nb1 > 0) {
            // third stream, beside the bulk update of step q: after the bulk
            // update of step q-1 (rows and columns Q) and P's panels
            if (q > 0) RGF_HIP(hipStreamWaitEvent(s3, evB[(q - 1) & 1], 0));
            RGF_HIP(hipStreamWaitEvent(s3, evS[q & 1], 0));
            // Rold = A[Q, J] - Aold[Q, :] A[P, J]; An[I, :] = A[I, Q] - Aold[I, :] A[P, Q]
            const GemmArgs<T> ga = from_source(gemm_args<T>(nb1, N, nb, Aold[q & 1].offset(p1, 0), A.offset(p0, 0),
                                                            Rold, cd(-1, 0), 1, Skip{0, 0, p0, p2}),
                                               A.offset(p1, 0));
            const GemmArgs<T> gb = from_source(gemm_args<T>(N, nb1, nb, Aold[q & 1], A.offset(p0, p1), An, cd(-1, 0), 1,
                                                            Skip{p0, p2, 0, 0}),
                                               A.offset(0, p1));
            if (nb2 > 0) {
                const GemmArgs<T> gc = gemm_args<T>(nb2, nb2, nb, Aold[q & 1].offset(p2, 0), A.offset(p0, p2),
                                                    A.offset(p2, p2), cd(-1, 0), 1, Skip{});
                cgemm_triple<T>(s3, ga, gb, gc, use3m, evT[q & 1]);
            } else {
                cgemm_pair<T>(s3, ga, gb, use3m, evT[q & 1]);
            }
        }
        if (timer) timer->begin(s2);
        {
            // whole skipped tile rows and columns left out of the grid when
            // the block edges are on the 64-wide tile grid
            const GemmArgs<T> g =
                gemm_args<T>(N, N, nb, Aold[q & 1], A.offset(p0, 0), A, cd(-1, 0), 1, Skip{p0, p2, p0, p2, p2, p3});
            constexpr int BK = GemmCfg<T>::BK2;
            bool done = false;
            done = launch_lean<T, true>(s2, g, use3m);
            if (!done)
                done = use3m ? launch_pair2<T, 64, 64, BK, 2, 2, false, 2, true, true>(s2, g, nullptr)
                             : launch_pair2<T, 64, 64, BK, 2, 2, false, 2, true, false>(s2, g, nullptr);
            if (!done)
                cgemm_skip<T>(s2, N, N, nb, Aold[q & 1], A.offset(p0, 0), A, cd(-1, 0), 1,
                              Skip{p0, p2, p0, p2, p2, p3}, use3m, nullptr, 0);
        }
        if (timer) {
            const double nn = (double)(N - nb - nb1);
            timer->end(s2, fl * (nn * nn - (double)nb2 * nb2) * nb);
        }
        RGF_HIP(hipEventRecord(evB[q & 1], s2));
        if (nb1 > 0) {
            RGF_HIP(hipStreamWaitEvent(s, evT[q & 1], 0));
            // Q's panels and step q+1 on (Q2, Q2) in one launch, then Q2's inverse
            panels_inplace(q + 1, evS[(q + 1) & 1]);
            if (nb2 > 0) launch_diag<T>(s, A, p2, nb2, D, dev_status);
        }
    }
    RGF_HIP(hipEventRecord(ov->ev[4], s2));
    RGF_HIP(hipStreamWaitEvent(s, ov->ev[4], 0));
    RGF_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ K1: slice assembly
template <typename T>
__global__ void assemble_dense_kernel(int64_t N, T* Mre, T* Mim, int64_t ld, const double* eps_re,
                                      const double* eps_im, double ecr, double eci, double k2r, double k2i,
                                      const T* Xre, const T* Xim, int64_t ldx, const T* Sre, const T* Sim,
                                      int64_t lds) {
    const int64_t total = N * N;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e / N, j = e % N;
        const int64_t o = i * ld + j;
        // the previous slice's matrix: in M itself, or in S (M is then only written)
        double vr, vi;
        if (Sre) {
            vr = -(double)Sre[i * lds + j];
            vi = -(double)Sim[i * lds + j];
        } else {
            vr = -(double)Mre[o];
            vi = -(double)Mim[o];
        }
        if (Xre) {
            vr -= (double)Xre[i * ldx + j];
            vi -= (double)Xim[i * ldx + j];
        }
        if (i == j) {
            const double er = eps_re ? eps_re[i] : ecr;
            const double ei = eps_re ? (eps_im ? eps_im[i] : 0.0) : eci;
            vr -= k2r * er - k2i * ei;
            vi -= k2r * ei + k2i * er;
        }
        Mre[o] = (T)vr;
        Mim[o] = (T)vi;
    }
}

template <typename T>
__global__ void csr_add_kernel(int32_t n, const int32_t* rowptr, const int32_t* col, const double* vre,
                               const double* vim, T* Mre, T* Mim, int64_t ld) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int q = rowptr[i]; q < rowptr[i + 1]; ++q) {
        const int64_t o = (int64_t)i * ld + col[q];
        Mre[o] = (T)((double)Mre[o] + vre[q]);
        Mim[o] = (T)((double)Mim[o] + vim[q]);
    }
}


template <typename T>
void assemble_slice(hipStream_t s, int64_t N, CMat<T> M, const double* eps_re, const double* eps_im,
                    cd eps_const, cd k2, const CMat<T>* Gextra, const DevCSR& A0, const CMat<T>* src) {
    hipLaunchKernelGGL((assemble_dense_kernel<T>), dim3(grid_for(N * N)), dim3(256), 0, s, N, M.re, M.im, M.ld,
                       eps_re, eps_im, eps_const.real(), eps_const.imag(), k2.real(), k2.imag(),
                       Gextra ? Gextra->re : (const T*)nullptr, Gextra ? Gextra->im : (const T*)nullptr,
                       Gextra ? Gextra->ld : (int64_t)0, src ? src->re : (const T*)nullptr,
                       src ? src->im : (const T*)nullptr, src ? src->ld : (int64_t)0);
    hipLaunchKernelGGL((csr_add_kernel<T>), dim3((A0.n + 255) / 256), dim3(256), 0, s, A0.n, A0.rowptr, A0.col,
                       A0.vre, A0.vim, M.re, M.im, M.ld);
    RGF_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ dense helpers
template <typename T>
__global__ void fill_kernel(int64_t rows, int64_t cols, T* re, T* im, int64_t ld, int ident) {
    const int64_t total = rows * cols;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e / cols, j = e % cols;
        re[i * ld + j] = (ident && i == j) ? T(1) : T(0);
        im[i * ld + j] = T(0);
    }
}

template <typename T>
void fill_zero(hipStream_t s, int64_t rows, int64_t cols, CMat<T> A) {
    if (rows <= 0 || cols <= 0) return;
    hipLaunchKernelGGL((fill_kernel<T>), dim3(grid_for(rows * cols)), dim3(256), 0, s, rows, cols, A.re, A.im,
                       A.ld, 0);
}

template <typename T>
void set_identity(hipStream_t s, int64_t N, CMat<T> A) {
    hipLaunchKernelGGL((fill_kernel<T>), dim3(grid_for(N * N)), dim3(256), 0, s, N, N, A.re, A.im, A.ld, 1);
}

// A synthetic slice matrix for timing the slice inverse: diagonal 4 (the
// stencil's), off-diagonal entries with independent real and imaginary parts
// uniform in [-a, a], a = 0.5 / sqrt(N) (spectral norm of the off-diagonal
// part about 0.6), from a hash of (seed, i, j).
template <typename T>
__global__ void tune_fill_kernel(int64_t N, T* re, T* im, int64_t ld, uint32_t seed) {
    const double a = 0.5 / sqrt((double)N);
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < N * N;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e / N, j = e % N;
        uint64_t h = (uint64_t)e * 0x9E3779B97F4A7C15ull + seed;
        h ^= h >> 31;
        h *= 0xBF58476D1CE4E5B9ull;
        h ^= h >> 29;
        const double u = (double)(h >> 40) * (1.0 / 16777216.0), v = (double)(h & 0xFFFFFF) * (1.0 / 16777216.0);
        re[i * ld + j] = (T)((2 * u - 1) * a + (i == j ? 4.0 : 0.0));
        im[i * ld + j] = (T)((2 * v - 1) * a);
    }
}

template <typename T>
void tune_fill(hipStream_t s, int64_t N, CMat<T> A, uint32_t seed) {
    if (N <= 0) return;
    hipLaunchKernelGGL((tune_fill_kernel<T>), dim3(grid_for(N * N)), dim3(256), 0, s, N, A.re, A.im, A.ld, seed);
    RGF_HIP(hipGetLastError());
}

template <typename Td, typename Ts>
__global__ void convert_kernel(int64_t rows, int64_t cols, const Ts* sre, const Ts* sim, int64_t lds, Td* dre,
                               Td* dim, int64_t ldd, double scale) {
    const int64_t total = rows * cols;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e / cols, j = e % cols;
        dre[i * ldd + j] = (Td)((double)sre[i * lds + j] * scale);
        dim[i * ldd + j] = (Td)((double)sim[i * lds + j] * scale);
    }
}

template <typename Td, typename Ts>
void convert(hipStream_t s, int64_t rows, int64_t cols, CMat<Ts> src, CMat<Td> dst, double scale) {
    if (rows <= 0 || cols <= 0) return;
    hipLaunchKernelGGL((convert_kernel<Td, Ts>), dim3(grid_for(rows * cols)), dim3(256), 0, s, rows, cols, src.re,
                       src.im, src.ld, dst.re, dst.im, dst.ld, scale);
}

template <typename T>
void copy_mat(hipStream_t s, int64_t rows, int64_t cols, CMat<T> src, CMat<T> dst) {
    convert<T, T>(s, rows, cols, src, dst, 1.0);
}

template <typename T>
__global__ void accumulate_kernel(int64_t rows, int64_t cols, const T* sre, const T* sim, int64_t lds,
                                  double* dre, double* dim, int64_t ldd, double scale) {
    const int64_t total = rows * cols;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e / cols, j = e % cols;
        dre[i * ldd + j] += (double)sre[i * lds + j] * scale;
        dim[i * ldd + j] += (double)sim[i * lds + j] * scale;
    }
}

template <typename T>
void accumulate(hipStream_t s, int64_t rows, int64_t cols, CMat<T> src, CMat<double> dst, double scale) {
    if (rows <= 0 || cols <= 0) return;
    hipLaunchKernelGGL((accumulate_kernel<T>), dim3(grid_for(rows * cols)), dim3(256), 0, s, rows, cols, src.re,
                       src.im, src.ld, dst.re, dst.im, dst.ld, scale);
}

template <typename T>
__global__ void add_kernel(int64_t rows, int64_t cols, const T* are, const T* aim, int64_t lda, const T* bre,
                           const T* bim, int64_t ldb, T* dre, T* dim, int64_t ldd) {
    const int64_t total = rows * cols;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e / cols, j = e % cols;
        dre[i * ldd + j] = are[i * lda + j] + bre[i * ldb + j];
        dim[i * ldd + j] = aim[i * lda + j] + bim[i * ldb + j];
    }
}

template <typename T>
void add_mat(hipStream_t s, int64_t rows, int64_t cols, CMat<T> a, CMat<T> b, CMat<T> dst) {
    if (rows <= 0 || cols <= 0) return;
    hipLaunchKernelGGL((add_kernel<T>), dim3(grid_for(rows * cols)), dim3(256), 0, s, rows, cols, a.re, a.im, a.ld,
                       b.re, b.im, b.ld, dst.re, dst.im, dst.ld);
}

// ------------------------------------------------------------------ K4: mode matrices
__global__ void fill_modes_kernel(int64_t n1, int64_t n2, const double* f1, const double* f2, const double* g,
                                  double* p1r, double* p1i, int64_t ld1, double* p2r, double* p2i, int64_t ld2) {
    const int64_t N = n1 * n2;
    const int64_t total = N * N;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e / N, a = e % N;
        const int64_t i1 = i % n1, i2 = i / n1, a1 = a % n1, a2 = a / n1;
        const double xr = f1[2 * (i1 * n1 + a1)], xi = f1[2 * (i1 * n1 + a1) + 1];
        const double yr = f2[2 * (i2 * n2 + a2)], yi = f2[2 * (i2 * n2 + a2) + 1];
        const double pr = xr * yr - xi * yi, pi = xr * yi + xi * yr;
        double gr = 1.0, gi = 0.0;
        if (g) {
            gr = g[2 * a];
            gi = g[2 * a + 1];
        }
        p1r[i * ld1 + a] = pr * gr - pi * gi;
        p1i[i * ld1 + a] = pr * gi + pi * gr;
        p2r[a * ld2 + i] = pr;
        p2i[a * ld2 + i] = -pi;
    }
}

void fill_modes(hipStream_t s, int64_t n1, int64_t n2, const double* f1, const double* f2, const double* g,
                CMat<double> P1, CMat<double> P2) {
    const int64_t N = n1 * n2;
    hipLaunchKernelGGL(fill_modes_kernel, dim3(grid_for(N * N)), dim3(256), 0, s, n1, n2, f1, f2, g, P1.re, P1.im,
                       P1.ld, P2.re, P2.im, P2.ld);
    RGF_HIP(hipGetLastError());
}

// Lead Green's function of a separable lead, G0 = phi diag(g) phi^H with
// phi = f1 (x) f2, in two contractions instead of an N^3 product:
//   M[a2][i1][j1] = sum_a1 f1[i1][a1] g[a1 + n1 a2] conj(f1[j1][a1])
//   G0[i][j]      = sum_a2 f2[i2][a2] M[a2][i1][j1] conj(f2[j2][a2])
// with i = i1 + n1 i2, j = j1 + n1 j2 (n1^3 n2 + N^2 n2 multiply-adds).
__global__ void lead_green_stage1_kernel(int64_t n1, int64_t n2, const double* f1, const double* g, double* M) {
    const int64_t total = n2 * n1 * n1;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j1 = e % n1, i1 = (e / n1) % n1, a2 = e / (n1 * n1);
        double sr = 0, si = 0;
        for (int64_t a1 = 0; a1 < n1; ++a1) {
            const double xr = f1[2 * (i1 * n1 + a1)], xi = f1[2 * (i1 * n1 + a1) + 1];
            const double yr = f1[2 * (j1 * n1 + a1)], yi = -f1[2 * (j1 * n1 + a1) + 1];
            const double gr = g[2 * (a1 + n1 * a2)], gi = g[2 * (a1 + n1 * a2) + 1];
            const double tr = xr * gr - xi * gi, ti = xr * gi + xi * gr;
            sr += tr * yr - ti * yi;
            si += tr * yi + ti * yr;
        }
        M[2 * e] = sr;
        M[2 * e + 1] = si;
    }
}

__global__ void lead_green_stage2_kernel(int64_t n1, int64_t n2, const double* f2, const double* M, double* Gre,
                                         double* Gim, int64_t ld) {
    const int64_t N = n1 * n2;
    for (int64_t i = blockIdx.y; i < N; i += gridDim.y) {
        const int64_t i1 = i % n1, i2 = i / n1;
        for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < N; j += (int64_t)gridDim.x * blockDim.x) {
            const int64_t j1 = j % n1, j2 = j / n1;
            double sr = 0, si = 0;
            for (int64_t a2 = 0; a2 < n2; ++a2) {
                const double xr = f2[2 * (i2 * n2 + a2)], xi = f2[2 * (i2 * n2 + a2) + 1];
                const double yr = f2[2 * (j2 * n2 + a2)], yi = -f2[2 * (j2 * n2 + a2) + 1];
                const double* m = M + 2 * ((a2 * n1 + i1) * n1 + j1);
                const double tr = xr * m[0] - xi * m[1], ti = xr * m[1] + xi * m[0];
                sr += tr * yr - ti * yi;
                si += tr * yi + ti * yr;
            }
            Gre[i * ld + j] = sr;
            Gim[i * ld + j] = si;
        }
    }
}

void lead_green_separable(hipStream_t s, int64_t n1, int64_t n2, const double* f1, const double* f2, const double* g,
                          double* M, CMat<double> G0) {
    const int64_t N = n1 * n2;
    if (N == 0) return;
    hipLaunchKernelGGL(lead_green_stage1_kernel, dim3(grid_for(n2 * n1 * n1)), dim3(256), 0, s, n1, n2, f1, g, M);
    const unsigned gx = (unsigned)std::min<int64_t>((N + 255) / 256, 64), gy = (unsigned)std::min<int64_t>(N, 65535);
    hipLaunchKernelGGL(lead_green_stage2_kernel, dim3(gx, gy), dim3(256), 0, s, n1, n2, f2, (const double*)M, G0.re,
                       G0.im, G0.ld);
    RGF_HIP(hipGetLastError());
}

// Selected mode columns: cols holds, per column q, (a1, a2, s_re, s_im) as
// doubles; phi_q(i) = f1[i1][a1] f2[i2][a2] with i = i1 + n1 i2.  rows = 0:
// D(i, q) = phi_q(i) s_q (N x n); rows = 1: D(q, i) = s_q conj(phi_q(i)) (n x N).
__global__ void fill_mode_cols_kernel(int64_t n1, int64_t n2, const double* f1, const double* f2, int64_t n,
                                      const double* cols, int rows, double* dr, double* di, int64_t ld) {
    const int64_t N = n1 * n2;
    const int64_t total = N * n;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t q = rows ? e / N : e % n, i = rows ? e % N : e / n;
        const int64_t i1 = i % n1, i2 = i / n1;
        const int64_t a1 = (int64_t)cols[4 * q], a2 = (int64_t)cols[4 * q + 1];
        const double sr = cols[4 * q + 2], si = cols[4 * q + 3];
        const double xr = f1[2 * (i1 * n1 + a1)], xi = f1[2 * (i1 * n1 + a1) + 1];
        const double yr = f2[2 * (i2 * n2 + a2)], yi = f2[2 * (i2 * n2 + a2) + 1];
        const double pr = xr * yr - xi * yi, pi = xr * yi + xi * yr;
        if (rows) {
            dr[q * ld + i] = sr * pr + si * pi;  // s conj(phi)
            di[q * ld + i] = si * pr - sr * pi;
        } else {
            dr[i * ld + q] = pr * sr - pi * si;
            di[i * ld + q] = pr * si + pi * sr;
        }
    }
}

void fill_mode_cols(hipStream_t s, int64_t n1, int64_t n2, const double* f1, const double* f2, int64_t n,
                    const double* cols, int rows, CMat<double> D) {
    const int64_t N = n1 * n2;
    if (N == 0 || n == 0) return;
    hipLaunchKernelGGL(fill_mode_cols_kernel, dim3(grid_for(N * n)), dim3(256), 0, s, n1, n2, f1, f2, n, cols, rows,
                       D.re, D.im, D.ld);
    RGF_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ residual (FP64)
// psi/r are (L+2) slices of N x m, row-major, slice stride N*m.
__global__ void residual_kernel(int64_t N, int64_t L, int64_t m, const double* pr, const double* pim, double* rr,
                                double* ri, const double* er, const double* ei, double eLr, double eLi, double eRr,
                                double eRi, double k2r, double k2i, const int32_t* rowptr, const int32_t* col,
                                const double* vre, const double* vim, const double* b0r, const double* b0i,
                                int64_t mL, const double* b1r, const double* b1i, int64_t mR, int64_t n0) {
    const int64_t S = N * m;
    const int64_t n = n0 + blockIdx.y;  // slice
    const unsigned um = (unsigned)m;
    for (unsigned rem = blockIdx.x * blockDim.x + threadIdx.x; rem < (unsigned)S; rem += gridDim.x * blockDim.x) {
        const unsigned i = rem / um, c = rem - i * um;
        const int64_t e = n * S + rem;
        // eps of slice n: leads at 0 and L+1, scattering slices 1..L
        double xr, xi;
        if (n == 0) {
            xr = eLr;
            xi = eLi;
        } else if (n == L + 1) {
            xr = eRr;
            xi = eRi;
        } else {
            xr = er[(n - 1) * N + i];
            xi = ei ? ei[(n - 1) * N + i] : 0.0;
        }
        // k2*eps
        const double ker = k2r * xr - k2i * xi, kei = k2r * xi + k2i * xr;
        const double psr = pr[e], psi = pim[e];
        // -(A0 psi)_i + k2 eps psi
        double accr = ker * psr - kei * psi;
        double acci = ker * psi + kei * psr;
        for (int q = rowptr[i]; q < rowptr[i + 1]; ++q) {
            const int64_t o = n * S + (int64_t)col[q] * m + c;
            const double ar = vre[q], ai = vim[q];
            accr -= ar * pr[o] - ai * pim[o];
            acci -= ar * pim[o] + ai * pr[o];
        }
        if (n > 0) {
            accr += pr[e - S];
            acci += pim[e - S];
        }
        if (n < L + 1) {
            accr += pr[e + S];
            acci += pim[e + S];
        }
        if (n == 0 && c < mL) {
            accr += b0r[i * mL + c];
            acci += b0i[i * mL + c];
        }
        if (n == L + 1 && c >= mL && c < mL + mR) {
            accr += b1r[i * mR + (c - mL)];
            acci += b1i[i * mR + (c - mL)];
        }
        rr[e] = accr;
        ri[e] = acci;
    }
}

void residual_stencil(hipStream_t s, int64_t N, int64_t L, int64_t m, const double* psi_re, const double* psi_im,
                      double* r_re, double* r_im, const double* eps_re, const double* eps_im, cd epsL, cd epsR,
                      cd k2, const DevCSR& A0, const double* b0_re, const double* b0_im, int64_t mL,
                      const double* b1_re, const double* b1_im, int64_t mR) {
    // grid: x over the N*m entries of a slice, y over up to 65535 slices
    if ((uint64_t)N * (uint64_t)m >= (1ull << 31)) throw Error(RGF_ERR_INVALID_ARGUMENT, "slice too large");
    for (int64_t n0 = 0; n0 < L + 2; n0 += 65535) {
        const dim3 grid((unsigned)grid_for(N * m, 256, 1024), (unsigned)std::min<int64_t>(65535, L + 2 - n0));
        hipLaunchKernelGGL(residual_kernel, grid, dim3(256), 0, s, N, L, m, psi_re, psi_im, r_re, r_im, eps_re,
                           eps_im, epsL.real(), epsL.imag(), epsR.real(), epsR.imag(), k2.real(), k2.imag(),
                           A0.rowptr, A0.col, A0.vre, A0.vim, b0_re, b0_im, mL, b1_re, b1_im, mR, n0);
    }
    RGF_HIP(hipGetLastError());
}

// Residual of the row system y H = c (the adjoint of H psi = b): y_n is
// m x N, r_n = c_n - y_n D_n + y_{n-1} + y_{n+1} with D_n = A0 - k2 diag(eps_n)
// (slices 0 and L+1: eps_L, eps_R; the caller adds the lead terms y_0 G0L
// and y_{L+1} G0R).  y_n D_n needs the columns of A0: AT is A0 transposed
// (CSR).  c is non-zero at slice L+1 only (m x N).
__global__ void residual_rows_kernel(int64_t N, int64_t L, int64_t m, const double* yr, const double* yim,
                                     double* rr, double* ri, const double* er, const double* ei, double eLr,
                                     double eLi, double eRr, double eRi, double k2r, double k2i,
                                     const int32_t* rowptr, const int32_t* col, const double* vre,
                                     const double* vim, const double* cr, const double* ci, int64_t n0) {
    const int64_t S = N * m;
    const int64_t n = n0 + blockIdx.y;
    const unsigned uN = (unsigned)N;
    for (unsigned rem = blockIdx.x * blockDim.x + threadIdx.x; rem < (unsigned)S; rem += gridDim.x * blockDim.x) {
        const unsigned j = rem / uN, i = rem - j * uN;
        const int64_t e = n * S + rem;
        double xr, xi;
        if (n == 0) {
            xr = eLr;
            xi = eLi;
        } else if (n == L + 1) {
            xr = eRr;
            xi = eRi;
        } else {
            xr = er[(n - 1) * N + i];
            xi = ei ? ei[(n - 1) * N + i] : 0.0;
        }
        const double ker = k2r * xr - k2i * xi, kei = k2r * xi + k2i * xr;
        const double yr0 = yr[e], yi0 = yim[e];
        double accr = ker * yr0 - kei * yi0;
        double acci = ker * yi0 + kei * yr0;
        for (int q = rowptr[i]; q < rowptr[i + 1]; ++q) {
            const int64_t o = n * S + (int64_t)j * N + col[q];
            const double ar = vre[q], ai = vim[q];
            accr -= ar * yr[o] - ai * yim[o];
            acci -= ar * yim[o] + ai * yr[o];
        }
        if (n > 0) {
            accr += yr[e - S];
            acci += yim[e - S];
        }
        if (n < L + 1) {
            accr += yr[e + S];
            acci += yim[e + S];
        }
        if (n == L + 1) {
            accr += cr[rem];
            acci += ci[rem];
        }
        rr[e] = accr;
        ri[e] = acci;
    }
}

void residual_rows(hipStream_t s, int64_t N, int64_t L, int64_t m, const double* y_re, const double* y_im,
                   double* r_re, double* r_im, const double* eps_re, const double* eps_im, cd epsL, cd epsR, cd k2,
                   const DevCSR& AT, const double* c_re, const double* c_im) {
    if ((uint64_t)N * (uint64_t)m >= (1ull << 31)) throw Error(RGF_ERR_INVALID_ARGUMENT, "slice too large");
    for (int64_t n0 = 0; n0 < L + 2; n0 += 65535) {
        const dim3 grid((unsigned)grid_for(N * m, 256, 1024), (unsigned)std::min<int64_t>(65535, L + 2 - n0));
        hipLaunchKernelGGL(residual_rows_kernel, grid, dim3(256), 0, s, N, L, m, y_re, y_im, r_re, r_im, eps_re,
                           eps_im, epsL.real(), epsL.imag(), epsR.real(), epsR.imag(), k2.real(), k2.imag(),
                           AT.rowptr, AT.col, AT.vre, AT.vim, c_re, c_im, n0);
    }
    RGF_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ field profiles
// out[(i + N * np) + N * (L + 2) * a] (interleaved complex) = psi at physical
// slice np, transverse site i, input a: internal slice (mirrored ? L+1-np : np),
// internal column col[a].
__global__ void pack_fields_kernel(int64_t N, int64_t L, int64_t m, int64_t nin, const double* pr, const double* pim,
                                   const int32_t* col, int mirrored, double* out) {
    const int64_t per = N * (L + 2);
    const int64_t total = per * nin;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t a = e / per, rem = e - a * per, np = rem / N, i = rem - np * N;
        const int64_t n = mirrored ? L + 1 - np : np;
        const int64_t src = (n * N + i) * m + col[a];
        out[2 * e] = pr[src];
        out[2 * e + 1] = pim[src];
    }
}

void pack_fields(hipStream_t s, int64_t N, int64_t L, int64_t m, int64_t nin, const double* psi_re,
                 const double* psi_im, const int32_t* col, int mirrored, double* out) {
    const int64_t total = N * (L + 2) * nin;
    if (total <= 0) return;
    hipLaunchKernelGGL(pack_fields_kernel, dim3(grid_for(total)), dim3(256), 0, s, N, L, m, nin, psi_re, psi_im, col,
                       mirrored, out);
    RGF_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ adjoint contraction
__global__ void grad_contract_kernel(int64_t N, int64_t m, const double* pr, const double* pim, int64_t ldp,
                                     const double* fr, const double* fi, int64_t ldf, double k2r, double k2i,
                                     double* gre, double* gim) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= N) return;
    double ar = 0, ai = 0;
    for (int64_t j = 0; j < m; ++j) {
        const double xr = pr[i * ldp + j], xi = pim[i * ldp + j];
        const double yr = fr[j * ldf + i], yi = fi[j * ldf + i];
        ar += xr * yr - xi * yi;
        ai += xr * yi + xi * yr;
    }
    const double ur = k2r * ar - k2i * ai, ui = k2r * ai + k2i * ar;
    gre[i] = 4.0 * ui;
    gim[i] = 4.0 * ur;
}

void grad_contract(hipStream_t s, int64_t N, int64_t m, CMat<double> psi, CMat<double> phi, cd k2, double* gre,
                   double* gim) {
    hipLaunchKernelGGL(grad_contract_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, N, m, psi.re, psi.im,
                       psi.ld, phi.re, phi.im, phi.ld, k2.real(), k2.imag(), gre, gim);
    RGF_HIP(hipGetLastError());
}

__global__ void max_abs_kernel(int64_t count, const double* re, const double* im, unsigned long long* out) {
    double m = 0.0;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < count;
         e += (int64_t)gridDim.x * blockDim.x) {
        const double a = re[e], b = im ? im[e] : 0.0;
        const double v = sqrt(a * a + b * b);
        if (!(v <= m)) m = v;  // NaN propagates
    }
    for (int off = 32; off >= 1; off >>= 1) {
        const double o = __shfl_xor(m, off);
        if (!(o <= m)) m = o;
    }
    if ((threadIdx.x & 63) == 0) {
        unsigned long long bits = isnan(m) ? 0x7ff8000000000000ULL : (unsigned long long)__double_as_longlong(m);
        atomicMax(out, bits);
    }
}

double max_abs(hipStream_t s, int64_t count, const double* re, const double* im, double* dev_scratch) {
    if (count <= 0) return 0.0;
    RGF_HIP(hipMemsetAsync(dev_scratch, 0, sizeof(double), s));
    hipLaunchKernelGGL(max_abs_kernel, dim3(grid_for(count, 256, 2048)), dim3(256), 0, s, count, re, im,
                       (unsigned long long*)dev_scratch);
    double h = 0;
    RGF_HIP(hipMemcpyAsync(&h, dev_scratch, sizeof(double), hipMemcpyDeviceToHost, s));
    RGF_HIP(hipStreamSynchronize(s));
    return h;
}

// ------------------------------------------------------------------ no-allocation workspace
// The device memory of rgf_dense_inverse_noalloc: 256 bytes of status words,
// then the slice-inverse workspace.  Each GPU holds its own copy at its own
// address (noalloc_scratch) until the process ends.
static __device__ __align__(256) unsigned char scratch[RGF_NOALLOC_SCRATCH_BYTES];

void* noalloc_scratch(int device) {
    static std::mutex mu;
    static std::map<int, void*> ptrs;
    std::lock_guard<std::mutex> lk(mu);
    const auto it = ptrs.find(device);
    if (it != ptrs.end()) return it->second;
    int cur = 0;
    RGF_HIP(hipGetDevice(&cur));
    if (cur != device) RGF_HIP(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    RGF_HIP(e);
    ptrs[device] = p;
    return p;
}

__global__ void zero_words_kernel(int* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

void zero_words(hipStream_t s, int* p, int n) {
    hipLaunchKernelGGL(zero_words_kernel, dim3(1), dim3(64), 0, s, p, n);
    RGF_HIP(hipGetLastError());
}

// planar (ld) <-> interleaved complex (re, im pairs, ld = N), row-major N x N
template <typename T>
__global__ void to_interleaved_kernel(int64_t N, const T* re, const T* im, int64_t ld, T* out) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < N * N; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e / N, j = e % N;
        out[2 * e] = re[i * ld + j];
        out[2 * e + 1] = im[i * ld + j];
    }
}
template <typename T>
__global__ void from_interleaved_kernel(int64_t N, const T* in, T* re, T* im, int64_t ld) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < N * N; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e / N, j = e % N;
        re[i * ld + j] = in[2 * e];
        im[i * ld + j] = in[2 * e + 1];
    }
}
template <typename T>
void to_interleaved(hipStream_t s, int64_t N, CMat<T> A, T* out) {
    hipLaunchKernelGGL((to_interleaved_kernel<T>), dim3(grid_for(N * N)), dim3(256), 0, s, N, A.re, A.im, A.ld, out);
    RGF_HIP(hipGetLastError());
}
template <typename T>
void from_interleaved(hipStream_t s, int64_t N, const T* in, CMat<T> A) {
    hipLaunchKernelGGL((from_interleaved_kernel<T>), dim3(grid_for(N * N)), dim3(256), 0, s, N, in, A.re, A.im, A.ld);
    RGF_HIP(hipGetLastError());
}

// bit 1 of the status word when either of two LAPACK-style info words is
// non-zero (an exactly zero pivot of the pivoted factorisation)
__global__ void fold_info_kernel(const int* info, int* status) {
    if (threadIdx.x == 0 && (info[0] != 0 || info[1] != 0)) atomicOr(status, 1);
}
void fold_info(hipStream_t s, const int* info, int* status) {
    hipLaunchKernelGGL(fold_info_kernel, dim3(1), dim3(64), 0, s, info, status);
    RGF_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ instantiations
#define RGF_INST(T)                                                                                          \
    template void cgemm<T>(hipStream_t, int64_t, int64_t, int64_t, CMat<T>, CMat<T>, CMat<T>, cd, int, int64_t, \
                           int64_t, int, T*, size_t);                                                         \
    template size_t inverse_workspace_elems<T>(int64_t, int);                                                 \
    template void invert_inplace<T>(hipStream_t, int64_t, CMat<T>, int, T*, int*, int, KernelTimer*,           \
                                    const Overlap*);                                                          \
    template void assemble_slice<T>(hipStream_t, int64_t, CMat<T>, const double*, const double*, cd, cd,      \
                                    const CMat<T>*, const DevCSR&, const CMat<T>*);                           \
    template void fill_zero<T>(hipStream_t, int64_t, int64_t, CMat<T>);                                       \
    template void set_identity<T>(hipStream_t, int64_t, CMat<T>);                                             \
    template void tune_fill<T>(hipStream_t, int64_t, CMat<T>, uint32_t);                                      \
    template void copy_mat<T>(hipStream_t, int64_t, int64_t, CMat<T>, CMat<T>);                               \
    template void accumulate<T>(hipStream_t, int64_t, int64_t, CMat<T>, CMat<double>, double);                \
    template void add_mat<T>(hipStream_t, int64_t, int64_t, CMat<T>, CMat<T>, CMat<T>);                      \
    template void to_interleaved<T>(hipStream_t, int64_t, CMat<T>, T*);                                       \
    template void from_interleaved<T>(hipStream_t, int64_t, const T*, CMat<T>);

RGF_INST(double)
RGF_INST(float)
template void convert<float, double>(hipStream_t, int64_t, int64_t, CMat<double>, CMat<float>, double);
template void convert<double, float>(hipStream_t, int64_t, int64_t, CMat<float>, CMat<double>, double);
template void convert<double, double>(hipStream_t, int64_t, int64_t, CMat<double>, CMat<double>, double);
template void convert<float, float>(hipStream_t, int64_t, int64_t, CMat<float>, CMat<float>, double);

}  // namespace rgf

// ==================================================================== batched engine
// Many independent mixed-precision solves in lockstep (BatchProb tables).
// Work maps hold (problem << 20) | local index.
namespace rgf {

__device__ __forceinline__ int map_prob(int32_t e) { return e >> 20; }
__device__ __forceinline__ int map_loc(int32_t e) { return e & 0xfffff; }

// GemmArgs of a plain product inside a kernel (every tile interior: the
// batched shapes are multiples of 64, rows 16-byte aligned)
template <typename T>
__device__ __forceinline__ GemmArgs<T> dev_args(int M, int N, int K, const T* Ar, const T* Ai, int64_t lda, const T* Br,
                                                const T* Bi, int64_t ldb, T* Cr, T* Ci, int64_t ldc, double alpha,
                                                int beta, int tiles_n) {
    GemmArgs<T> g;
    g.Are = Ar;
    g.Aim = Ai;
    g.lda = lda;
    g.Bre = Br;
    g.Bim = Bi;
    g.ldb = ldb;
    g.Cre = Cr;
    g.Cim = Ci;
    g.ldc = ldc;
    g.Xre = Cr;
    g.Xim = Ci;
    g.ldx = ldc;
    g.Yre = nullptr;
    g.Yim = nullptr;
    g.ldy = 0;
    g.y_rlo = g.y_rhi = g.y_clo = g.y_chi = 0;
    g.M = M;
    g.N = N;
    g.K = K;
    g.alpha_re = alpha;
    g.alpha_im = 0.0;
    g.beta = beta;
    g.srlo = g.srhi = g.sclo = g.schi = 0;
    g.sblo = g.sbhi = 0;
    g.vec_ok = 1;
    g.nz = 1;
    g.kc = K;
    g.zstride = 0;
    g.tiles_n = tiles_n;
    g.tiles = 0;
    g.cr_lo = g.cc_lo = INT32_MAX;
    g.cr_n = g.cc_n = 0;
    g.amax = nullptr;
    g.ctr = nullptr;
    g.ctr_clo = g.ctr_chi = 0;
    g.ctr_wait = 0;
    g.fail = nullptr;
    return g;
}

// Panel workspace of a problem: D (64 x 64), R (64 x NP), C (NP x 64), planar.
__device__ __forceinline__ float* ws_D(const BatchProb& pb) { return pb.ws; }
__device__ __forceinline__ float* ws_R(const BatchProb& pb) { return pb.ws + 2 * 64 * 64; }
__device__ __forceinline__ float* ws_C(const BatchProb& pb) { return pb.ws + 2 * 64 * 64 + 2 * 64 * (size_t)pb.NP; }

// g_{n-1}(i, j) as it stands once the write-back of block [p0w, p0w + 64) is
// done: entries with exactly one index in the block come from the panel
// workspace (R for the block's rows, C for its columns) while `pending`.
__device__ __forceinline__ void prev_entry(const BatchProb& pb, const float* Pr, const float* Pi, bool pending,
                                           int p0w, int i, int j, float& vr, float& vi) {
    const int NP = pb.NP;
    const bool ri = i >= p0w && i < p0w + 64, cj = j >= p0w && j < p0w + 64;
    if (pending && ri != cj) {
        if (ri) {
            const float* R = ws_R(pb);
            vr = R[(size_t)(i - p0w) * NP + j];
            vi = R[(size_t)64 * NP + (size_t)(i - p0w) * NP + j];
        } else {
            const float* C = ws_C(pb);
            vr = C[(size_t)i * 64 + (j - p0w)];
            vi = C[(size_t)NP * 64 + (size_t)i * 64 + (j - p0w)];
        }
    } else {
        vr = Pr[(size_t)i * NP + j];
        vi = Pi[(size_t)i * NP + j];
    }
}

// Entry (i, j) of the padded slice matrix of slice n (1..L+1) given
// g_{n-1}(i, j): A0 - k^2 diag(eps_n) - g_{n-1}, the right lead's
// A0 - k^2 eps_R - G0R - g_L at n = L + 1, the identity on the padding;
// summed in FP64, rounded once.
__device__ __forceinline__ void slice_entry(const BatchProb& pb, int n, int i, int j, float gr, float gi, float& outr,
                                            float& outi) {
    if (i >= pb.N || j >= pb.N) {
        outr = (i == j) ? 1.f : 0.f;
        outi = 0.f;
        return;
    }
    double vr = -(double)gr, vi = -(double)gi;
    const size_t NP = pb.NP;
    if (n == pb.L + 1) {
        const size_t o = (size_t)i * NP + j;
        vr -= pb.G0R[o];
        vi -= pb.G0R[NP * NP + o];
    }
    if (i == j) {
        double er, ei;
        if (n <= pb.L) {
            er = pb.eps_re[(size_t)(n - 1) * pb.N + i];
            ei = pb.eps_im ? pb.eps_im[(size_t)(n - 1) * pb.N + i] : 0.0;
        } else {
            er = pb.eRr;
            ei = pb.eRi;
        }
        vr -= pb.k2r * er - pb.k2i * ei;
        vi -= pb.k2r * ei + pb.k2i * er;
    }
    for (int q = pb.rowptr[i]; q < pb.rowptr[i + 1]; ++q)
        if (pb.col[q] == j) {
            vr += pb.vre[q];
            vi += pb.vim[q];
        }
    outr = (float)vr;
    outi = (float)vi;
}

// Rows of a slice handled by one work group of the row-wise passes (four
// wavefronts, four rows each: a wavefront takes a row, its lanes the columns).
constexpr int kRowsPerGroup = kBatchRows;
static_assert(kRowsPerGroup == 4, "a wavefront per row, four per work group");

// Round r, diagonal blocks: work group w inverts the diagonal block of
// problem w's step r (slice n = 1 + r / nb, block q = r % nb); at q = 0 it
// first assembles that block of the slice matrix (for nb > 1 the block lies
// outside the pending write-back's rows and columns; for nb = 1 nothing is
// pending).  Runs beside batch_rows_kernel of the same round.
__global__ __launch_bounds__(256) void batch_diag_kernel(const BatchProb* P, const int* rbase, int k) {
    const int r = *rbase + k;
    const BatchProb& pb = P[blockIdx.x];
    const int nb = pb.nb, NP = pb.NP;
    if (r >= (pb.L + 1) * nb) return;
    const int n = 1 + r / nb, q = r % nb;
    const size_t NN = (size_t)NP * NP;
    float* Fr = pb.F + (size_t)n * 2 * NN;
    float* Fi = Fr + NN;
    if (q == 0) {
        const float* Pr = Fr - 2 * NN;
        const float* Pi = Pr + NN;
        for (int e = (int)threadIdx.x; e < 64 * 64; e += 256) {
            const int i = e >> 6, j = e & 63;
            float gr, gi, vr, vi;
            prev_entry(pb, Pr, Pi, false, 0, i, j, gr, gi);
            slice_entry(pb, n, i, j, gr, gi, vr, vi);
            Fr[(size_t)i * NP + j] = vr;
            Fi[(size_t)i * NP + j] = vi;
        }
        __syncthreads();
    }
    gj_diag_blk<float, 64>(Fr, Fi, NP, q * 64, 64, ws_D(pb), ws_D(pb) + 64 * 64, 64, pb.status);
}

// Round r, row pass (beside the diagonal blocks): the pending write-back of
// step r - 1 (its block's rows from R, its block's columns from C) and, when
// step r starts a slice (q = 0), the assembly of that slice, which folds in
// the write-back of the previous slice's last block.  The diagonal group owns
// block (0, 0) of an assembled slice.  Step (L + 1) nb only writes back.
// A wavefront per row, four consecutive columns per lane (16-byte loads and
// stores; a lane's four columns lie in one 64-wide block).
// map: (problem, group of kRowsPerGroup rows).
__global__ __launch_bounds__(256) void batch_rows_kernel(const BatchProb* P, const int* rbase, int k,
                                                         const int32_t* map) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    const int r = *rbase + k;
    const int32_t e0 = map[blockIdx.x];
    const BatchProb& pb = P[map_prob(e0)];
    const int nb = pb.nb, NP = pb.NP, N = pb.N, T = (pb.L + 1) * nb;
    if (r > T) return;
    const bool asmb = r < T && (r % nb) == 0;
    const bool pending = r >= 1 && nb > 1;
    if (!asmb && !pending) return;
    const size_t NN = (size_t)NP * NP;
    const int ps = r - 1;
    const int p0w = pending ? (ps % nb) * 64 : 0;
    float* Wr = pb.F + (size_t)(pending ? 1 + ps / nb : 0) * 2 * NN;  // the pending step's slice
    float* Wi = Wr + NN;
    const float* R = ws_R(pb);
    const float* C = ws_C(pb);
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6), lane = (int)threadIdx.x & 63;
    const int i = map_loc(e0) * kRowsPerGroup + wave;
    if (i >= NP) return;
    const bool rowP = pending && i >= p0w && i < p0w + 64;
    if (!asmb) {
        if (rowP) {
            const size_t ro = (size_t)(i - p0w) * NP;
            for (int j = 4 * lane; j < NP; j += 256) {
                if (j >= p0w && j < p0w + 64) continue;
                *(f4*)(Wr + (size_t)i * NP + j) = *(const f4*)(R + ro + j);
                *(f4*)(Wi + (size_t)i * NP + j) = *(const f4*)(R + (size_t)64 * NP + ro + j);
            }
        } else {
            Wr[(size_t)i * NP + p0w + lane] = C[(size_t)i * 64 + lane];
            Wi[(size_t)i * NP + p0w + lane] = C[(size_t)NP * 64 + (size_t)i * 64 + lane];
        }
        return;
    }
    // row i of slice n: A_n(i, :) - g_{n-1}(i, :)
    const int n = 1 + r / nb;
    float* Fr = pb.F + (size_t)n * 2 * NN;
    float* Fi = Fr + NN;
    const float* Pr = Fr - 2 * NN;
    const float* Pi = Pr + NN;
    const bool real_row = i < N;
    double dr = 0.0, di = 0.0;  // -k^2 eps_i on the diagonal
    int rp0 = 0, rp1 = 0;
    if (real_row) {
        double er, ei;
        if (n <= pb.L) {
            er = pb.eps_re[(size_t)(n - 1) * N + i];
            ei = pb.eps_im ? pb.eps_im[(size_t)(n - 1) * N + i] : 0.0;
        } else {
            er = pb.eRr;
            ei = pb.eRi;
        }
        dr = -(pb.k2r * er - pb.k2i * ei);
        di = -(pb.k2r * ei + pb.k2i * er);
        rp0 = pb.rowptr[i];
        rp1 = pb.rowptr[i + 1];
    }
    const bool lead = n == pb.L + 1;
    for (int j = 4 * lane; j < NP; j += 256) {
        if (i < 64 && j < 64) continue;
        const bool colP = j >= p0w && j < p0w + 64;
        f4 gr, gi;
        if (pending && rowP && !colP) {
            const size_t ro = (size_t)(i - p0w) * NP + j;
            gr = *(const f4*)(R + ro);
            gi = *(const f4*)(R + (size_t)64 * NP + ro);
        } else if (pending && !rowP && colP) {
            const size_t co = (size_t)i * 64 + (j - p0w);
            gr = *(const f4*)(C + co);
            gi = *(const f4*)(C + (size_t)NP * 64 + co);
        } else {
            gr = *(const f4*)(Pr + (size_t)i * NP + j);
            gi = *(const f4*)(Pi + (size_t)i * NP + j);
        }
        if (pending && rowP != colP) {
            *(f4*)(Wr + (size_t)i * NP + j) = gr;
            *(f4*)(Wi + (size_t)i * NP + j) = gi;
        }
        f4 vr, vi;
        if (!real_row || j >= N) {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                vr[u] = (i == j + u) ? 1.f : 0.f;
                vi[u] = 0.f;
            }
        } else {
            double ar[4], ai[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                ar[u] = -(double)gr[u];
                ai[u] = -(double)gi[u];
            }
            if (lead) {
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const size_t o = (size_t)i * NP + j + u;
                    ar[u] -= pb.G0R[o];
                    ai[u] -= pb.G0R[NN + o];
                }
            }
            if (i >= j && i < j + 4) {
                ar[i - j] += dr;
                ai[i - j] += di;
            }
            for (int q = rp0; q < rp1; ++q) {
                const int cq = pb.col[q] - j;
                if (cq >= 0 && cq < 4) {
                    ar[cq] += pb.vre[q];
                    ai[cq] += pb.vim[q];
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                // columns j + u >= N (the padding) of a real row are zero
                vr[u] = (j + u < N) ? (float)ar[u] : 0.f;
                vi[u] = (j + u < N) ? (float)ai[u] : 0.f;
            }
        }
        *(f4*)(Fr + (size_t)i * NP + j) = vr;
        *(f4*)(Fi + (size_t)i * NP + j) = vi;
    }
}

// Round r, second launch: R = D A[P, :] (columns P skipped) and
// C = -A[:, P] D (rows P skipped; its largest entry is the step's largest
// multiplier) of every problem's step r.  Tiles [0, nb) of a problem are R's,
// [nb, 2 nb) C's.
template <bool M3>
__global__ __launch_bounds__(256) void batch_k2_kernel(const BatchProb* P, const int* rbase, int k,
                                                       const int32_t* map2) {
    const int r = *rbase + k;
    const int32_t e0 = map2[blockIdx.x];
    const BatchProb& pb = P[map_prob(e0)];
    const int t = map_loc(e0);
    const int nb = pb.nb, NP = pb.NP;
    if (r >= (pb.L + 1) * nb) return;
    const int n = 1 + r / nb, p0 = (r % nb) * 64;
    const size_t NN = (size_t)NP * NP;
    float* Fr = pb.F + (size_t)n * 2 * NN;
    float* Fi = Fr + NN;
    float* D = ws_D(pb);
    float* R = ws_R(pb);
    float* C = ws_C(pb);
    if (t < nb) {
        GemmArgs<float> g = dev_args<float>(64, NP, 64, D, D + 64 * 64, 64, Fr + (size_t)p0 * NP,
                                            Fi + (size_t)p0 * NP, NP, R, R + (size_t)64 * NP, NP, 1.0, 0, nb);
        g.sclo = p0;
        g.schi = p0 + 64;
        cgemm2_tile<float, 64, 64, 16, 2, 2, false, 0, true, M3, true>(g, t);
    } else {
        GemmArgs<float> g = dev_args<float>(NP, 64, 64, Fr + p0, Fi + p0, NP, D, D + 64 * 64, 64, C,
                                            C + (size_t)NP * 64, 64, -1.0, 0, 1);
        g.srlo = p0;
        g.srhi = p0 + 64;
        g.amax = (unsigned*)(pb.status + 1);
        cgemm2_tile<float, 64, 64, 16, 2, 2, false, 0, true, M3, true>(g, t - nb);
    }
}

// Round r, third launch: the trailing update A[I, J] -= A[I, P] R[P, J] of
// every problem's step r over the (nb - 1)^2 tiles outside block row and
// column P (A[:, P] still holds the old column: it is written back in the
// next round's first launch).
template <bool M3>
__global__ __launch_bounds__(256, 4) void batch_k3_kernel(const BatchProb* P, const int* rbase, int k,
                                                          const int32_t* map3) {
    const int r = *rbase + k;
    const int32_t e0 = map3[blockIdx.x];
    const BatchProb& pb = P[map_prob(e0)];
    const int t = map_loc(e0);
    const int nb = pb.nb, NP = pb.NP;
    if (r >= (pb.L + 1) * nb) return;
    const int n = 1 + r / nb, q = r % nb, p0 = q * 64;
    const int ti0 = t / (nb - 1), tj0 = t % (nb - 1);
    const int ti = ti0 + (ti0 >= q ? 1 : 0), tj = tj0 + (tj0 >= q ? 1 : 0);
    const size_t NN = (size_t)NP * NP;
    float* Fr = pb.F + (size_t)n * 2 * NN;
    float* Fi = Fr + NN;
    float* R = ws_R(pb);
    const GemmArgs<float> g = dev_args<float>(NP, NP, 64, Fr + p0, Fi + p0, NP, R, R + (size_t)64 * NP, NP, Fr, Fi,
                                              NP, -1.0, 1, nb);
    cgemm2_tile<float, 64, 64, 16, 2, 2, false, 0, true, M3, false>(g, ti * nb + tj);
}

__global__ void batch_advance_kernel(int* rbase, int k) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *rbase += k;
}

// F[0] = (float) G0L, the identity on the padding
__global__ __launch_bounds__(256) void batch_first_slice_kernel(const BatchProb* P, const int32_t* map) {
    const int32_t e0 = map[blockIdx.x];
    const BatchProb& pb = P[map_prob(e0)];
    const int NP = pb.NP, N = pb.N;
    const int64_t NN = (int64_t)NP * NP;
    const int64_t lo = (int64_t)map_loc(e0) * kBatchChunk, hi = min(lo + (int64_t)kBatchChunk, NN);
    for (int64_t e = lo + threadIdx.x; e < hi; e += 256) {
        const int i = (int)(e / NP), j = (int)(e % NP);
        float vr = (i == j) ? 1.f : 0.f, vi = 0.f;
        if (i < N && j < N) {
            vr = (float)pb.G0L[e];
            vi = (float)pb.G0L[NN + e];
        }
        pb.F[e] = vr;
        pb.F[NN + e] = vi;
    }
}

// ------------------------------------------------------------------ batched refinement
// Residual r = b - H psi at slice n, transverse site i (< N), column c:
// b holds B_L at slice 0 (columns < mL) and B_R at slice L + 1 (columns
// [mL, mL + mR)); H psi = (A0 - k^2 eps_n) psi_n - psi_{n-1} - psi_{n+1},
// with the lead terms G0L psi_0 and G0R psi_{L+1} (W) at the end slices.
// zero_psi: psi = 0 (the first correction), r = b.  The per-row values
// (k^2 eps_n(i) and the stencil row) come from ResRow.
struct ResRow {
    double ker, kei;
    int rp0, rp1;
};

__device__ __forceinline__ ResRow res_row(const BatchProb& pb, int n, int i) {
    const int S = pb.L + 2;
    double er, ei;
    if (n == 0) {
        er = pb.eLr;
        ei = pb.eLi;
    } else if (n == S - 1) {
        er = pb.eRr;
        ei = pb.eRi;
    } else {
        er = pb.eps_re[(size_t)(n - 1) * pb.N + i];
        ei = pb.eps_im ? pb.eps_im[(size_t)(n - 1) * pb.N + i] : 0.0;
    }
    ResRow rw;
    rw.ker = pb.k2r * er - pb.k2i * ei;
    rw.kei = pb.k2r * ei + pb.k2i * er;
    rw.rp0 = pb.rowptr[i];
    rw.rp1 = pb.rowptr[i + 1];
    return rw;
}

__device__ __forceinline__ void batch_residual(const BatchProb& pb, const ResRow& rw, int n, int i, int c,
                                               bool zero_psi, double& rr, double& ri) {
    const int m = pb.m, S = pb.L + 2;
    const size_t SE = (size_t)pb.NP * m;
    const size_t o = (size_t)i * m + c;
    rr = 0.0;
    ri = 0.0;
    if ((n == 0 && c < pb.mL) || (n == S - 1 && c >= pb.mL && c < pb.mL + pb.mR)) {
        rr = pb.B[o];
        ri = pb.B[SE + o];
    }
    if (zero_psi) return;
    const double* pr = pb.psi + (size_t)n * SE;
    const double* pim = pb.psi + (size_t)S * SE + (size_t)n * SE;
    const double psr = pr[o], psi = pim[o];
    double ar = rw.ker * psr - rw.kei * psi, ai = rw.ker * psi + rw.kei * psr;
    for (int q = rw.rp0; q < rw.rp1; ++q) {
        const size_t oq = (size_t)pb.col[q] * m + c;
        const double vr = pb.vre[q], vi = pb.vim[q];
        ar -= vr * pr[oq] - vi * pim[oq];
        ai -= vr * pim[oq] + vi * pr[oq];
    }
    if (n > 0) {
        ar += pr[o - SE];
        ai += pim[o - SE];
    }
    if (n < S - 1) {
        ar += pr[o + SE];
        ai += pim[o + SE];
    }
    if (n == 0) {
        ar += pb.W[o];
        ai += pb.W[SE + o];
    }
    if (n == S - 1) {
        ar += pb.W[2 * SE + o];
        ai += pb.W[3 * SE + o];
    }
    rr += ar;
    ri += ai;
}

// Residual pass: r = b - H psi over every slice (first: psi = 0, only the
// lead slices, where b lives; the other slices of z are zero), stored in
// FP32 into z (the previous correction is already in psi; z_n holds r_n
// until the forward sweep replaces it by g_n (r_n + z_{n-1})), and max |r|
// per problem.  A wavefront per row, kResRowsPerWave rows each; map:
// (problem, group of 4 kResRowsPerWave rows of the S N (first: 2 N) rows).
constexpr int kResRowsPerWave = kBatchResRows / 4;
__global__ __launch_bounds__(256) void batch_resmax_kernel(const BatchProb* P, const int32_t* map, int first) {
    const int32_t e0 = map[blockIdx.x];
    const BatchProb& pb = P[map_prob(e0)];
    const int N = pb.N, S = pb.L + 2, m = pb.m;
    const int rows = first ? 2 * N : S * N;
    const size_t SE = (size_t)pb.NP * m;
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6), lane = (int)threadIdx.x & 63;
    double mx = 0.0;
    for (int rr = 0; rr < kResRowsPerWave; ++rr) {
        const int u = (map_loc(e0) * 4 + wave) * kResRowsPerWave + rr;
        if (u >= rows) break;
        const int sl = u / N, i = u - sl * N;
        const int n = first ? (sl ? S - 1 : 0) : sl;
        const ResRow rw = res_row(pb, n, i);
        float* zr = pb.z + (size_t)n * SE + (size_t)i * m;
        float* zi = pb.z + (size_t)S * SE + (size_t)n * SE + (size_t)i * m;
        for (int c = lane; c < m; c += 64) {
            double rr_, ri_;
            batch_residual(pb, rw, n, i, c, first != 0, rr_, ri_);
            zr[c] = (float)rr_;
            zi[c] = (float)ri_;
            const double v = sqrt(rr_ * rr_ + ri_ * ri_);
            if (!(v <= mx)) mx = v;  // NaN propagates
        }
    }
    for (int off = 32; off >= 1; off >>= 1) {
        const double o = __shfl_xor(mx, off);
        if (!(o <= mx)) mx = o;
    }
    if (lane == 0) {
        const unsigned long long bits =
            isnan(mx) ? 0x7ff8000000000000ULL : (unsigned long long)__double_as_longlong(mx);
        atomicMax(pb.rmax, bits);
    }
}

// forward step n: t = r_n + z_{n-1} (r_n in z_n, z_{n-1} = g_{n-1} t_{n-1}
// already), over the N x m real rows (the padding rows stay zero); map:
// (problem, chunk of kBatchChunk elements of N m, both planes), 16-byte loads.
__global__ __launch_bounds__(256) void batch_fwd_rhs_kernel(const BatchProb* P, const int32_t* map, int n) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    const int32_t e0 = map[blockIdx.x];
    const BatchProb& pb = P[map_prob(e0)];
    const int S = pb.L + 2;
    if (n >= S) return;
    const int tot = pb.N * pb.m;  // a multiple of 4 (m is)
    const size_t SE = (size_t)pb.NP * pb.m;
    const float* rr = pb.z + (size_t)n * SE;
    const float* ri = pb.z + (size_t)S * SE + (size_t)n * SE;
    const int lo = map_loc(e0) * kBatchChunk, hi = min(lo + kBatchChunk, tot);
    for (int e = lo + 4 * (int)threadIdx.x; e < hi; e += 1024) {
        f4 vr = *(const f4*)(rr + e), vi = *(const f4*)(ri + e);
        if (n > 0) {
            vr += *(const f4*)(rr - SE + e);
            vi += *(const f4*)(ri - SE + e);
        }
        *(f4*)(pb.t + e) = vr;
        *(f4*)(pb.t + SE + e) = vi;
    }
}

// z_n = g_n t (forward) or z_n += g_n z_{n+1} with n = L - k (backward):
// NP x m products, K = NP, tiles NP/64 x ceil(m/64).  first (forward, the
// first correction): before the last slice only the left inputs' columns
// [0, ceil4(mL)) are non-zero (the right-hand side is B_L at slice 0), so the
// products run on those columns; z starts zeroed.
template <bool M3, bool BWD>
__global__ __launch_bounds__(256, 4) void batch_sweep_gemm_kernel(const BatchProb* P, const int32_t* map, int step,
                                                                  int first) {
    const int32_t e0 = map[blockIdx.x];
    const BatchProb& pb = P[map_prob(e0)];
    const int S = pb.L + 2, NP = pb.NP;
    const int n = BWD ? S - 2 - step : step;
    if (n < 0 || n >= S) return;
    int m = pb.m;
    const int tn = (m + 63) / 64;
    if (!BWD && first && n < S - 1) {
        m = min(m, (pb.mL + 3) / 4 * 4);
        if ((map_loc(e0) % tn) * 64 >= m) return;
    }
    const size_t NN = (size_t)NP * NP, SE = (size_t)NP * pb.m;
    const float* Fr = pb.F + (size_t)n * 2 * NN;
    const float* Fi = Fr + NN;
    float* zr = pb.z + (size_t)n * SE;
    float* zi = pb.z + (size_t)S * SE + (size_t)n * SE;
    const float* br = BWD ? zr + SE : pb.t;
    const float* bi = BWD ? zi + SE : pb.t + SE;
    const GemmArgs<float> g =
        dev_args<float>(NP, m, NP, Fr, Fi, NP, br, bi, pb.m, zr, zi, pb.m, 1.0, BWD ? 1 : 0, tn);
    cgemm2_tile<float, 64, 64, 16, 2, 2, false, 0, true, M3, false>(g, map_loc(e0));
}

// W0 = G0L psi_0, W1 = G0R psi_{L+1} (FP64); tiles [0, h) W0's, [h, 2h) W1's
__global__ __launch_bounds__(256) void batch_lead_kernel(const BatchProb* P, const int32_t* map) {
    const int32_t e0 = map[blockIdx.x];
    const BatchProb& pb = P[map_prob(e0)];
    const int S = pb.L + 2, NP = pb.NP, m = pb.m;
    const int tn = (m + 63) / 64, h = (NP / 64) * tn;
    int t = map_loc(e0);
    const bool right = t >= h;
    if (right) t -= h;
    const size_t NN = (size_t)NP * NP, SE = (size_t)NP * m;
    const double* G = right ? pb.G0R : pb.G0L;
    const size_t so = right ? (size_t)(S - 1) * SE : 0;
    double* W = pb.W + (right ? 2 * SE : 0);
    const GemmArgs<double> g = dev_args<double>(NP, m, NP, G, G + NN, NP, pb.psi + so, pb.psi + (size_t)S * SE + so, m,
                                                W, W + SE, m, 1.0, 0, tn);
    cgemm2_tile<double, 64, 64, 16, 2, 2, false, 0, true, false, false>(g, t);
}

// T = [C_L psi_0; C_R psi_{L+1}] (FP64); tiles [0, hL) the left block's
__global__ __launch_bounds__(256) void batch_project_kernel(const BatchProb* P, const int32_t* map) {
    const int32_t e0 = map[blockIdx.x];
    const BatchProb& pb = P[map_prob(e0)];
    const int S = pb.L + 2, NP = pb.NP, m = pb.m;
    const int tn = (m + 63) / 64, hL = ((pb.npo + 63) / 64) * tn;
    int t = map_loc(e0);
    const bool right = t >= hL;
    if (right) t -= hL;
    const int rows = right ? pb.nqo : pb.npo;
    if (rows == 0) return;
    const size_t SE = (size_t)NP * m, TE = (size_t)(pb.npo + pb.nqo) * m;
    const double* Cm = right ? pb.CR : pb.CL;
    const size_t so = right ? (size_t)(S - 1) * SE : 0;
    double* T = pb.T + (right ? (size_t)pb.npo * m : 0);
    const GemmArgs<double> g = dev_args<double>(rows, m, NP, Cm, Cm + (size_t)rows * NP, NP, pb.psi + so,
                                                pb.psi + (size_t)S * SE + so, m, T, T + TE, m, 1.0, 0, tn);
    cgemm2_tile<double, 64, 64, 16, 2, 2, false, 0, true, false, false>(g, t);
}

// psi += z over every slice (first: psi = z, psi was never written)
__global__ __launch_bounds__(256) void batch_accumulate_kernel(const BatchProb* P, const int32_t* map, int first) {
    const int32_t e0 = map[blockIdx.x];
    const BatchProb& pb = P[map_prob(e0)];
    const int64_t tot = (int64_t)(pb.L + 2) * pb.NP * pb.m;
    const int64_t lo = (int64_t)map_loc(e0) * kBatchChunk, hi = min(lo + (int64_t)kBatchChunk, tot);
    for (int64_t e = lo + threadIdx.x; e < hi; e += 256) {
        const double a = (double)pb.z[e], b = (double)pb.z[tot + e];
        pb.psi[e] = first ? a : pb.psi[e] + a;
        pb.psi[tot + e] = first ? b : pb.psi[tot + e] + b;
    }
}

// ------------------------------------------------------------------ launchers
void batch_round(hipStream_t s, hipStream_t s2, hipEvent_t fork, hipEvent_t join, const BatchProb* P, int np,
                 const int* rbase, int k, const int32_t* k1map, int n1, const int32_t* k2map, int n2,
                 const int32_t* k3map, int n3, int use3m) {
    // the row pass on the second stream, beside the diagonal blocks
    RGF_HIP(hipEventRecord(fork, s));
    RGF_HIP(hipStreamWaitEvent(s2, fork, 0));
    if (n1 > 0) hipLaunchKernelGGL(batch_rows_kernel, dim3((unsigned)n1), dim3(256), 0, s2, P, rbase, k, k1map);
    RGF_HIP(hipEventRecord(join, s2));
    hipLaunchKernelGGL(batch_diag_kernel, dim3((unsigned)np), dim3(256), 0, s, P, rbase, k);
    RGF_HIP(hipStreamWaitEvent(s, join, 0));
    if (n2 > 0) {
        if (use3m)
            hipLaunchKernelGGL((batch_k2_kernel<true>), dim3((unsigned)n2), dim3(256), 0, s, P, rbase, k, k2map);
        else
            hipLaunchKernelGGL((batch_k2_kernel<false>), dim3((unsigned)n2), dim3(256), 0, s, P, rbase, k, k2map);
    }
    if (n3 > 0) {
        if (use3m)
            hipLaunchKernelGGL((batch_k3_kernel<true>), dim3((unsigned)n3), dim3(256), 0, s, P, rbase, k, k3map);
        else
            hipLaunchKernelGGL((batch_k3_kernel<false>), dim3((unsigned)n3), dim3(256), 0, s, P, rbase, k, k3map);
    }
    RGF_HIP(hipGetLastError());
}

void batch_advance(hipStream_t s, int* rbase, int k) {
    hipLaunchKernelGGL(batch_advance_kernel, dim3(1), dim3(64), 0, s, rbase, k);
}

void batch_first_slice(hipStream_t s, const BatchProb* P, const int32_t* map, int nchunks) {
    if (nchunks <= 0) return;
    hipLaunchKernelGGL(batch_first_slice_kernel, dim3((unsigned)nchunks), dim3(256), 0, s, P, map);
    RGF_HIP(hipGetLastError());
}

void batch_resmax(hipStream_t s, const BatchProb* P, const int32_t* map, int nchunks, int first) {
    if (nchunks <= 0) return;
    hipLaunchKernelGGL(batch_resmax_kernel, dim3((unsigned)nchunks), dim3(256), 0, s, P, map, first);
    RGF_HIP(hipGetLastError());
}

void batch_lead_terms(hipStream_t s, const BatchProb* P, const int32_t* map, int ntiles) {
    if (ntiles <= 0) return;
    hipLaunchKernelGGL(batch_lead_kernel, dim3((unsigned)ntiles), dim3(256), 0, s, P, map);
    RGF_HIP(hipGetLastError());
}

void batch_fwd_rhs(hipStream_t s, const BatchProb* P, const int32_t* map, int nchunks, int n) {
    if (nchunks <= 0) return;
    hipLaunchKernelGGL(batch_fwd_rhs_kernel, dim3((unsigned)nchunks), dim3(256), 0, s, P, map, n);
}

void batch_fwd_gemm(hipStream_t s, const BatchProb* P, const int32_t* map, int ntiles, int n, int use3m, int first) {
    if (ntiles <= 0) return;
    if (use3m)
        hipLaunchKernelGGL((batch_sweep_gemm_kernel<true, false>), dim3((unsigned)ntiles), dim3(256), 0, s, P, map, n,
                           first);
    else
        hipLaunchKernelGGL((batch_sweep_gemm_kernel<false, false>), dim3((unsigned)ntiles), dim3(256), 0, s, P, map,
                           n, first);
}

void batch_bwd_gemm(hipStream_t s, const BatchProb* P, const int32_t* map, int ntiles, int k, int use3m) {
    if (ntiles <= 0) return;
    if (use3m)
        hipLaunchKernelGGL((batch_sweep_gemm_kernel<true, true>), dim3((unsigned)ntiles), dim3(256), 0, s, P, map, k,
                           0);
    else
        hipLaunchKernelGGL((batch_sweep_gemm_kernel<false, true>), dim3((unsigned)ntiles), dim3(256), 0, s, P, map, k,
                           0);
}

void batch_accumulate(hipStream_t s, const BatchProb* P, const int32_t* map, int nchunks, int first) {
    if (nchunks <= 0) return;
    hipLaunchKernelGGL(batch_accumulate_kernel, dim3((unsigned)nchunks), dim3(256), 0, s, P, map, first);
    RGF_HIP(hipGetLastError());
}

void batch_project(hipStream_t s, const BatchProb* P, const int32_t* map, int ntiles) {
    if (ntiles <= 0) return;
    hipLaunchKernelGGL(batch_project_kernel, dim3((unsigned)ntiles), dim3(256), 0, s, P, map);
    RGF_HIP(hipGetLastError());
}

}  // namespace rgf
