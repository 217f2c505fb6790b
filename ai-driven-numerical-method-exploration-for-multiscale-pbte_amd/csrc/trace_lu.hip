// Device sparse direct solve of the GSIS-HDG trace system K u = f.
//
// Replaces the reference's SuperLU factorisation and solve of the global
// trace matrix (solver-python/pybte/acceleration/global_solve.py:120-124
// factorise, :190-195 solve) with a nested-dissection multifrontal solver
// that runs on the device:
//
// * symbolic phase (host, once per mesh): the cells are split by recursive
//   coordinate bisection of their centroids into a binary tree; a face
//   belongs to the lowest tree node whose cells hold both elements it
//   couples (the separator faces of that node; interior faces of a leaf);
//   a node's front is its own faces followed by the faces of its ancestors
//   that touch its cells.  Every block K[f, g] is assembled into the front of
//   the deeper of the owners of f and g.
// * numeric phase (device, level by level from the leaves): the fronts of a
//   level are assembled (original blocks, then the update matrices of child
//   0, then child 1: a fixed summation order), and the fully summed block of
//   each front is eliminated by the blocked Gauss-Jordan sweep with row
//   interchanges among its fully summed rows (partial pivoting).  The sweep
//   leaves [F11^-1 P^T, F11^-1 F12] in the fully summed rows, -F21 F11^-1 P^T
//   below them and the Schur complement F22 - F21 F11^-1 F12 (the update
//   matrix passed to the parent) in the rest.
// * solve (device): forward from the leaves, c = b2 - W P b1 passed to the
//   parent; backward from the root, x1 = [Z V] [P b1; -x2].  Every step is a
//   dense matrix-vector product over a front's stored rows: no triangular
//   recurrences, and each row is reduced in a fixed order (the result does
//   not depend on the launch).
//
// All arithmetic is binary64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "pbte_hip.h"

namespace {

#define TL_TRY(expr)                                                        \
    do {                                                                    \
        hipError_t e_ = (expr);                                             \
        if (e_ != hipSuccess) {                                             \
            throw std::string(#expr " failed: ") + hipGetErrorString(e_);  \
        }                                                                   \
    } while (0)

constexpr int KB = 32;        // pivot columns per sweep step

// stored factor rows are padded to a multiple of 4 values (zeros), so that
// every row starts 16-byte aligned in binary32 and binary64 and the solve
// reads it in 16-byte words
__host__ __device__ constexpr int row_ld(int n) { return (n + 3) & ~3; }
constexpr int TILE = 128;     // update tile (rows and columns)

struct Front {                // one tree node, device and host
    int64_t ws;               // front matrix offset in its level's workspace (doubles)
    int64_t zv;               // [Z V] rows (nfs x n) in the factor storage
    int64_t w;                // W rows (ncb x nfs) in the factor storage
    int64_t pan;              // panel scratch (n x KB), row copy (KB x n), block inverse (KB x KB)
    int32_t n, nfs, ncb;
    int32_t faces;            // offset of the front's faces (own, then ancestors')
    int32_t piv;              // offset of nfs pivot rows / forward values y
    int32_t cvec;             // offset of the ncb forward contributions
    int32_t child[2];         // node ids, -1 when absent
    int32_t cmap[2];          // offsets of the child face -> front position maps
    int32_t gsrc;             // offset of the n per-row child sources of the forward gather
};

struct Level {
    std::vector<int32_t> nodes;
    int64_t ws_size = 0, pan_size = 0;
    int32_t n_max = 0, nfs_max = 0, ncb_max = 0, ncb_child_max = 0;
    int32_t node_off = 0;             // into the device node list
    int32_t rec_off = 0, rec_n = 0;   // original-block records
    int32_t fw_off = 0, fw_n = 0;     // forward work items (front, row0)
    int32_t bw_off = 0, bw_n = 0;     // backward work items
    int32_t gi_off = 0, gi_n = 0;     // forward gather items (front, row0), 256 rows each
    int32_t lpr_fw = 64, lpr_bw = 64; // lanes per row of the solve's products
    bool has_children = false;
};

// ---------------------------------------------------------------------------
// device kernels: numeric factorisation
// ---------------------------------------------------------------------------
__global__ void k_fill(double* __restrict__ p, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

__global__ void k_narrow(const double* __restrict__ in, float* __restrict__ out, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (float)in[i];
}

__global__ void k_set_int(int32_t* __restrict__ p, int32_t v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// rec = {block, node, row position (scalars), column position (scalars)}
__global__ void k_assemble(const int4* __restrict__ rec, int nrec, const double* __restrict__ val, int b,
                           const Front* __restrict__ fr, double* __restrict__ ws)
{
    const int64_t bb = (int64_t)b * b;
    const int64_t total = (int64_t)nrec * bb;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
         g += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = g / bb;
        const int e = (int)(g - r * bb);
        const int i = e / b, j = e - (e / b) * b;
        const int4 q = rec[r];
        const Front f = fr[q.y];
        ws[f.ws + (int64_t)(q.z + i) * f.n + q.w + j] = val[(int64_t)q.x * bb + e];
    }
}

// adds the update matrix of child `slot` of every front of the level:
// blockIdx.x = front, blockIdx.y = chunk of 8 rows of the child's matrix
__global__ void k_extend(const int32_t* __restrict__ nodes, int slot, const Front* __restrict__ fr,
                         const int32_t* __restrict__ fmap, int b, double* __restrict__ ws,
                         const double* __restrict__ wsc)
{
    const Front v = fr[nodes[blockIdx.x]];
    const int c = v.child[slot];
    if (c < 0) return;
    const Front ch = fr[c];
    const double* S = wsc + ch.ws + (int64_t)ch.nfs * ch.n + ch.nfs;
    const int32_t* map = fmap + v.cmap[slot];
    double* M = ws + v.ws;
    const int i0 = blockIdx.y * 8;
    for (int i = i0; i < min(i0 + 8, ch.ncb); ++i) {
        const int pi = map[i / b] * b + i % b;
        for (int j = threadIdx.x; j < ch.ncb; j += blockDim.x) {
            const int pj = map[j / b] * b + j % b;
            M[(int64_t)pi * v.n + pj] += S[(int64_t)i * ch.n + j];
        }
    }
}

// Panel of KB pivot columns [j0, j0 + kbe) of each front: copies the columns
// (all n rows) to the panel scratch, sweeps them pivot by pivot with the
// pivot row chosen among the remaining fully summed rows (largest
// magnitude, first on ties), and records the interchanges.
__global__ void __launch_bounds__(1024) k_panel(const int32_t* __restrict__ nodes, const Front* __restrict__ fr,
                                                const double* __restrict__ ws, double* __restrict__ pan, int j0,
                                                int32_t* __restrict__ ipiv, int32_t* __restrict__ perm,
                                                int32_t* __restrict__ status)
{
    __shared__ double s_val[1024];
    __shared__ int s_idx[1024];
    const Front f = fr[nodes[blockIdx.x]];
    if (j0 >= f.nfs) return;
    const int n = f.n, kbe = min(KB, f.nfs - j0), tid = threadIdx.x, nt = blockDim.x;
    const double* M = ws + f.ws;
    double* P = pan + f.pan;
    if (j0 == 0)
        for (int k = tid; k < f.nfs; k += nt) perm[f.piv + k] = k;
    for (int64_t e = tid; e < (int64_t)n * KB; e += nt) {
        const int i = (int)(e / KB), c = (int)(e - (int64_t)i * KB);
        P[e] = c < kbe ? M[(int64_t)i * n + j0 + c] : 0.0;
    }
    __syncthreads();
    for (int k = 0; k < kbe; ++k) {
        const int r0 = j0 + k;
        double best = -1.0;
        int bi = r0;
        for (int i = r0 + tid; i < f.nfs; i += nt) {
            const double a = fabs(P[(int64_t)i * KB + k]);
            if (a > best) { best = a; bi = i; }
        }
        s_val[tid] = best;
        s_idx[tid] = bi;
        __syncthreads();
        for (int s = nt / 2; s > 0; s >>= 1) {
            if (tid < s) {
                const double o = s_val[tid + s];
                const int oi = s_idx[tid + s];
                if (o > s_val[tid] || (o == s_val[tid] && oi < s_idx[tid])) {
                    s_val[tid] = o;
                    s_idx[tid] = oi;
                }
            }
            __syncthreads();
        }
        const int p = s_idx[0];
        const bool singular = !(s_val[0] > 0.0);
        __syncthreads();
        if (singular) {
            // no pivot: the rest of the panel is left unswept, with no row
            // interchanges recorded (the later kernels read them)
            for (int q = k + tid; q < kbe; q += nt) ipiv[f.piv + j0 + q] = j0 + q;
            if (tid == 0) atomicMin(status, f.faces);   // a front's first face id as the tag
            return;
        }
        if (p != r0) {
            for (int c = tid; c < KB; c += nt) {
                const double t = P[(int64_t)r0 * KB + c];
                P[(int64_t)r0 * KB + c] = P[(int64_t)p * KB + c];
                P[(int64_t)p * KB + c] = t;
            }
        }
        if (tid == 0) {
            ipiv[f.piv + r0] = p;
            const int t = perm[f.piv + r0];
            perm[f.piv + r0] = perm[f.piv + p];
            perm[f.piv + p] = t;
        }
        __syncthreads();
        const double inv = 1.0 / P[(int64_t)r0 * KB + k];
        for (int64_t e = tid; e < (int64_t)n * kbe; e += nt) {
            const int i = (int)(e / kbe), c = (int)(e - (int64_t)i * kbe);
            if (i == r0 || c == k) continue;
            const double t = P[(int64_t)i * KB + k] * inv;
            P[(int64_t)i * KB + c] = fma(-t, P[(int64_t)r0 * KB + c], P[(int64_t)i * KB + c]);
        }
        __syncthreads();
        for (int c = tid; c < kbe; c += nt)
            if (c != k) P[(int64_t)r0 * KB + c] *= inv;
        for (int i = tid; i < n; i += nt)
            if (i != r0) P[(int64_t)i * KB + k] = -P[(int64_t)i * KB + k] * inv;
        if (tid == 0) P[(int64_t)r0 * KB + k] = inv;
        __syncthreads();
    }
}

// Diagonal-block pivoting (the fast path): the KB x KB diagonal block of
// the step, inverted in LDS by Gauss-Jordan with interchanges among the
// block's own rows, stored as the exact inverse (interchanges folded back
// into its columns), so that no row of the front moves.  One workgroup per
// front; a block without a nonzero pivot sets the status.
__global__ void __launch_bounds__(256) k_block_inv(const int32_t* __restrict__ nodes, const Front* __restrict__ fr,
                                                   const double* __restrict__ ws, double* __restrict__ pan, int j0,
                                                   int32_t* __restrict__ perm, int32_t* __restrict__ status)
{
    __shared__ double A[KB][KB + 1];
    __shared__ int sp[KB];
    __shared__ int s_p;
    __shared__ double s_v;
    const Front f = fr[nodes[blockIdx.x]];
    if (j0 >= f.nfs) return;
    const int kbe = min(KB, f.nfs - j0), tid = threadIdx.x, n = f.n;
    const double* M = ws + f.ws;
    if (j0 == 0)
        for (int k = tid; k < f.nfs; k += blockDim.x) perm[f.piv + k] = k;
    for (int e = tid; e < KB * KB; e += blockDim.x) {
        const int r = e / KB, c = e - (e / KB) * KB;
        A[r][c] = (r < kbe && c < kbe) ? M[(int64_t)(j0 + r) * n + j0 + c] : (r == c ? 1.0 : 0.0);
    }
    if (tid < KB) sp[tid] = tid;
    __syncthreads();
    for (int k = 0; k < kbe; ++k) {
        if (tid < 64) {
            // pivot among rows k..kbe-1 of the block: one wave, butterfly on (|a|, row)
            double v = -1.0;
            int r = k;
            if (tid >= k && tid < kbe) {
                v = fabs(A[tid][k]);
                r = tid;
            }
#pragma unroll
            for (int m = 32; m > 0; m >>= 1) {
                const double ov = __shfl_xor(v, m, 64);
                const int orr = __shfl_xor(r, m, 64);
                if (ov > v || (ov == v && orr < r)) {
                    v = ov;
                    r = orr;
                }
            }
            if (tid == 0) {
                s_p = r;
                s_v = v;
            }
        }
        __syncthreads();
        const int p = s_p;
        if (!(s_v > 0.0)) {
            if (tid == 0) atomicMin(status, f.faces);
            return;
        }
        if (p != k) {
            for (int c = tid; c < KB; c += blockDim.x) {
                const double t = A[k][c];
                A[k][c] = A[p][c];
                A[p][c] = t;
            }
            if (tid == 0) {
                const int t = sp[k];
                sp[k] = sp[p];
                sp[p] = t;
            }
        }
        __syncthreads();
        const double inv = 1.0 / A[k][k];
        for (int e = tid; e < KB * KB; e += blockDim.x) {
            const int r = e / KB, c = e - (e / KB) * KB;
            if (r == k || c == k) continue;
            A[r][c] = fma(-(A[r][k] * inv), A[k][c], A[r][c]);
        }
        __syncthreads();
        if (tid < KB) {
            if (tid != k) {
                A[k][tid] *= inv;
                A[tid][k] = -A[tid][k] * inv;
            } else {
                A[k][k] = inv;
            }
        }
        __syncthreads();
    }
    // inverse of the block: Ainv[r][sp[c]] = A[r][c]
    double* Ai = pan + f.pan + 2 * (int64_t)n * KB;
    for (int e = tid; e < KB * KB; e += blockDim.x) {
        const int r = e / KB, c = e - (e / KB) * KB;
        Ai[r * KB + sp[c]] = A[r][c];
    }
}

// The panel of the step from the block inverse: pivot rows P[J] = Ainv,
// every other row P[i] = -M[i][J] Ainv (the sweep of the panel columns on
// the block).  blockIdx.x = front, blockIdx.y = chunk of 64 rows.
__global__ void __launch_bounds__(256) k_panel_apply(const int32_t* __restrict__ nodes, const Front* __restrict__ fr,
                                                     const double* __restrict__ ws, double* __restrict__ pan, int j0)
{
    __shared__ double Ai[KB][KB];
    const Front f = fr[nodes[blockIdx.x]];
    if (j0 >= f.nfs) return;
    const int n = f.n, kbe = min(KB, f.nfs - j0), tid = threadIdx.x;
    const int i0 = blockIdx.y * 64;
    if (i0 >= n) return;
    const double* M = ws + f.ws;
    double* P = pan + f.pan;
    const double* Ag = pan + f.pan + 2 * (int64_t)n * KB;
    for (int e = tid; e < KB * KB; e += blockDim.x) Ai[e / KB][e % KB] = Ag[e];
    __syncthreads();
    // thread: one row, 8 of the KB columns (4 threads per row... 64 rows x 4)
    const int r = i0 + (tid >> 2), cq = (tid & 3) * (KB / 4);
    if (r >= n) return;
    double out[KB / 4];
    if (r >= j0 && r < j0 + kbe) {
#pragma unroll
        for (int c = 0; c < KB / 4; ++c) out[c] = Ai[r - j0][cq + c];
    } else {
#pragma unroll
        for (int c = 0; c < KB / 4; ++c) out[c] = 0.0;
        const double* row = M + (int64_t)r * n + j0;
        for (int k = 0; k < kbe; ++k) {
            const double x = row[k];
#pragma unroll
            for (int c = 0; c < KB / 4; ++c) out[c] = fma(-x, Ai[k][cq + c], out[c]);
        }
    }
#pragma unroll
    for (int c = 0; c < KB / 4; ++c) P[(int64_t)r * KB + cq + c] = cq + c < kbe ? out[c] : 0.0;
}

// The pivot rows (outside the panel; zero inside it) copied to the row
// scratch, no interchanges.  blockIdx.x = front, blockIdx.y = 256 columns.
__global__ void k_copy_rows(const int32_t* __restrict__ nodes, const Front* __restrict__ fr,
                            const double* __restrict__ ws, double* __restrict__ pan, int j0)
{
    const Front f = fr[nodes[blockIdx.x]];
    if (j0 >= f.nfs) return;
    const int c = blockIdx.y * blockDim.x + threadIdx.x;
    if (c >= f.n) return;
    const int kbe = min(KB, f.nfs - j0), n = f.n;
    const double* M = ws + f.ws;
    double* T = pan + f.pan + (int64_t)n * KB;
    const bool in_panel = c >= j0 && c < j0 + kbe;
    for (int k = 0; k < KB; ++k)
        T[(int64_t)k * n + c] = (in_panel || k >= kbe) ? 0.0 : M[(int64_t)(j0 + k) * n + c];
}

// Residual of the factorisation on a test vector: y = K x - b per row, and
// the row sums of |K| (block rows of the pattern)
__global__ void k_bsr_residual(const int32_t* __restrict__ ptr, const int32_t* __restrict__ col,
                               const double* __restrict__ val, int F, int b, const double* __restrict__ x,
                               const double* __restrict__ rhs, double* __restrict__ res, double* __restrict__ rowabs)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (int64_t)F * b) return;
    const int f = (int)(g / b), i = (int)(g - (int64_t)f * b);
    double s = 0.0, a = 0.0;
    for (int e = ptr[f]; e < ptr[f + 1]; ++e) {
        const double* blk = val + ((int64_t)e * b + i) * b;
        const double* xv = x + (int64_t)col[e] * b;
        for (int j = 0; j < b; ++j) {
            s = fma(blk[j], xv[j], s);
            a += fabs(blk[j]);
        }
    }
    res[g] = s - rhs[g];
    rowabs[g] = a;
}

// Row interchanges of the step on the columns outside the panel, then the
// pivot rows (outside the panel; zero inside it) copied to the row scratch.
// blockIdx.x = front, blockIdx.y = chunk of 256 columns.
__global__ void k_swap_rows(const int32_t* __restrict__ nodes, const Front* __restrict__ fr,
                            double* __restrict__ ws, double* __restrict__ pan, int j0,
                            const int32_t* __restrict__ ipiv)
{
    const Front f = fr[nodes[blockIdx.x]];
    if (j0 >= f.nfs) return;
    const int c = blockIdx.y * blockDim.x + threadIdx.x;
    if (c >= f.n) return;
    const int kbe = min(KB, f.nfs - j0), n = f.n;
    double* M = ws + f.ws;
    double* T = pan + f.pan + (int64_t)n * KB;
    const bool in_panel = c >= j0 && c < j0 + kbe;
    if (!in_panel) {
        for (int k = 0; k < kbe; ++k) {
            const int r0 = j0 + k, p = ipiv[f.piv + r0];
            if (p != r0 && p > r0 && p < f.nfs) {
                const double t = M[(int64_t)r0 * n + c];
                M[(int64_t)r0 * n + c] = M[(int64_t)p * n + c];
                M[(int64_t)p * n + c] = t;
            }
        }
    }
    for (int k = 0; k < KB; ++k)
        T[(int64_t)k * n + c] = (in_panel || k >= kbe) ? 0.0 : M[(int64_t)(j0 + k) * n + c];
}

// The rank-KB update of the sweep step on 128 x 128 tiles of every front:
// outside the panel columns M[i][c] = [i not a pivot row] M[i][c]
// + sum_k P[i][k] T[k][c]; inside them, the swept panel is stored.
// 256 threads, 8 x 8 values each (rows ty*8 + a, columns tx*8 + q), the
// panel and row tiles staged in LDS in chunks of KC columns / rows (the row
// tile column-interleaved, so that the lanes of a wave read consecutive
// words).  blockIdx.x = tile (row-major over the largest front of the
// level), blockIdx.y = front.
constexpr int KC = 16;
__global__ void __launch_bounds__(256) k_update(const int32_t* __restrict__ nodes, const Front* __restrict__ fr,
                                                double* __restrict__ ws, const double* __restrict__ pan, int j0,
                                                int tiles_per_row)
{
    __shared__ double sP[TILE][KC + 1];
    __shared__ double sT[KC][TILE];
    const Front f = fr[nodes[blockIdx.y]];
    if (j0 >= f.nfs) return;
    const int n = f.n;
    const int nt = (n + TILE - 1) / TILE;
    const int ti = blockIdx.x / tiles_per_row, tj = blockIdx.x - (blockIdx.x / tiles_per_row) * tiles_per_row;
    if (ti >= nt || tj >= nt) return;
    const int kbe = min(KB, f.nfs - j0);
    const int r0 = ti * TILE, c0 = tj * TILE;
    const double* P = pan + f.pan;
    const double* T = P + (int64_t)n * KB;
    double* M = ws + f.ws;
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    double acc[8][8];
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int q = 0; q < 8; ++q) acc[a][q] = 0.0;
    for (int k0 = 0; k0 < kbe; k0 += KC) {
        __syncthreads();
        for (int e = tid; e < TILE * KC; e += 256) {
            const int i = e / KC, k = e - (e / KC) * KC;
            sP[i][k] = (r0 + i < n) ? P[(int64_t)(r0 + i) * KB + k0 + k] : 0.0;
        }
        for (int e = tid; e < TILE * KC; e += 256) {
            const int k = e / TILE, c = e - (e / TILE) * TILE;       // column c0 + c of row k0 + k
            sT[k][(c & 7) * 16 + (c >> 3)] = (c0 + c < n) ? T[(int64_t)(k0 + k) * n + c0 + c] : 0.0;
        }
        __syncthreads();
#pragma unroll 4
        for (int k = 0; k < KC; ++k) {
            double pa[8], tb[8];
#pragma unroll
            for (int a = 0; a < 8; ++a) pa[a] = sP[ty * 8 + a][k];
#pragma unroll
            for (int q = 0; q < 8; ++q) tb[q] = sT[k][q * 16 + tx];
#pragma unroll
            for (int a = 0; a < 8; ++a)
#pragma unroll
                for (int q = 0; q < 8; ++q) acc[a][q] = fma(pa[a], tb[q], acc[a][q]);
        }
    }
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const int i = r0 + ty * 8 + a;
        if (i >= n) continue;
        const bool pivot_row = i >= j0 && i < j0 + kbe;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int c = c0 + tx * 8 + q;
            if (c >= n) continue;
            double* m = M + (int64_t)i * n + c;
            if (c >= j0 && c < j0 + kbe)
                *m = P[(int64_t)i * KB + (c - j0)];
            else
                *m = (pivot_row ? 0.0 : *m) + acc[a][q];
        }
    }
}

// factors out of the front matrix: [Z V] rows and W = -(rows below);
// blockIdx.x = front, blockIdx.y = chunk of 4 rows
__global__ void k_extract(const int32_t* __restrict__ nodes, const Front* __restrict__ fr,
                          const double* __restrict__ ws, double* __restrict__ fac)
{
    const Front f = fr[nodes[blockIdx.x]];
    const double* M = ws + f.ws;
    const int ldz = row_ld(f.n), ldw = row_ld(f.nfs);
    for (int i = blockIdx.y * 4; i < min(blockIdx.y * 4 + 4, f.n); ++i) {
        if (i < f.nfs) {
            for (int j = threadIdx.x; j < ldz; j += blockDim.x)
                fac[f.zv + (int64_t)i * ldz + j] = j < f.n ? M[(int64_t)i * f.n + j] : 0.0;
        } else {
            for (int j = threadIdx.x; j < ldw; j += blockDim.x)
                fac[f.w + (int64_t)(i - f.nfs) * ldw + j] = j < f.nfs ? -M[(int64_t)i * f.n + j] : 0.0;
        }
    }
}

// ---------------------------------------------------------------------------
// device kernels: solve
// ---------------------------------------------------------------------------

// forward assembly of the fronts of a level, one thread per front row k:
// the row j it takes (j = perm[k] for the fully summed rows, whose values
// go to y in pivot order; j = k for the contribution rows, which start the
// front's contributions c) is b(j) from the right-hand side (0 for an
// ancestor's row), plus the first child's contribution to it, plus the
// second's
__global__ void __launch_bounds__(256) k_fwd_gather(const int2* __restrict__ items, const Front* __restrict__ fr,
                                                    const int32_t* __restrict__ faces, const int2* __restrict__ gsrc,
                                                    const int32_t* __restrict__ perm, int b,
                                                    const double* __restrict__ rhs, double* __restrict__ y,
                                                    double* __restrict__ cv)
{
    const int2 it = items[blockIdx.x];
    const Front f = fr[it.x];
    const int k = it.y + (int)threadIdx.x;
    if (k >= f.n) return;
    const int j = k < f.nfs ? perm[f.piv + k] : k;
    double v = j < f.nfs ? rhs[(int64_t)faces[f.faces + j / b] * b + j % b] : 0.0;
    const int2 g = gsrc[f.gsrc + j];
    if (g.x >= 0) v += cv[g.x];
    if (g.y >= 0) v += cv[g.y];
    if (k < f.nfs)
        y[f.piv + k] = v;
    else
        cv[f.cvec + k - f.nfs] = v;
}

// sum over the LPR lanes of a row group (butterfly within the group)
template <int LPR>
__device__ inline double group_sum(double v)
{
#pragma unroll
    for (int m = LPR / 2; m > 0; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

template <typename TF> struct Wide;
template <> struct Wide<double> { using type = double2; static constexpr int n = 2; };
template <> struct Wide<float> { using type = float4; static constexpr int n = 4; };

// dot product of a stored row (ld values, 16-byte aligned, zero past the
// row's own length) with a vector of ld values by the LPR lanes of a group:
// each lane reads 16-byte words k = V (lane + LPR j), V = 16 / sizeof(TF),
// into two accumulators (even, odd j), sums in binary64
template <int LPR, typename TF>
__device__ inline double row_dot(const TF* __restrict__ row, const double* __restrict__ v, int ld, int lane)
{
    constexpr int V = Wide<TF>::n;
    using W = typename Wide<TF>::type;
    double s0 = 0.0, s1 = 0.0;
    int k = lane * V;
    for (; k + LPR * V < ld; k += 2 * LPR * V) {
        const W a = *reinterpret_cast<const W*>(row + k);
        const W c = *reinterpret_cast<const W*>(row + k + LPR * V);
        const TF* ea = reinterpret_cast<const TF*>(&a);
        const TF* ec = reinterpret_cast<const TF*>(&c);
#pragma unroll
        for (int j = 0; j < V; ++j) {
            s0 = fma((double)ea[j], v[k + j], s0);
            s1 = fma((double)ec[j], v[k + LPR * V + j], s1);
        }
    }
    if (k < ld) {
        const W a = *reinterpret_cast<const W*>(row + k);
        const TF* ea = reinterpret_cast<const TF*>(&a);
#pragma unroll
        for (int j = 0; j < V; ++j) s0 = fma((double)ea[j], v[k + j], s0);
    }
    return group_sum<LPR>(s0 + s1);
}

// c[i] -= W[i] . y over the rows of a work item (rows it.y .. it.y + R):
// 256 / LPR row groups of LPR lanes; the factors in TF (binary64, or their
// binary32 copy), the sums in binary64
template <int LPR, typename TF>
__global__ void __launch_bounds__(256) k_fwd_gemv(const int2* __restrict__ items, const Front* __restrict__ fr,
                                                  const TF* __restrict__ fac, const double* __restrict__ y,
                                                  double* __restrict__ cv, int R)
{
    constexpr int G = 256 / LPR;
    const int2 it = items[blockIdx.x];
    const Front f = fr[it.x];
    const int lane = threadIdx.x % LPR, grp = threadIdx.x / LPR;
    const double* yv = y + f.piv;
    for (int r = it.y + grp; r < min(it.y + R, f.ncb); r += G) {
        const double s = row_dot<LPR, TF>(fac + f.w + (int64_t)r * row_ld(f.nfs), yv, row_ld(f.nfs), lane);
        if (lane == 0) cv[f.cvec + r] -= s;
    }
}

// x[own faces] = [Z V] [y; -x(ancestors' faces)] over the rows of a work
// item (as k_fwd_gemv); the input vector is staged in LDS
template <int LPR, typename TF>
__global__ void __launch_bounds__(256) k_bwd(const int2* __restrict__ items, const Front* __restrict__ fr,
                                             const int32_t* __restrict__ faces, int b,
                                             const TF* __restrict__ fac, const double* __restrict__ y,
                                             double* __restrict__ x, int R)
{
    extern __shared__ double u[];
    constexpr int G = 256 / LPR;
    const int2 it = items[blockIdx.x];
    const Front f = fr[it.x];
    const int ld = row_ld(f.n);
    for (int j = threadIdx.x; j < ld; j += blockDim.x)
        u[j] = j < f.nfs ? y[f.piv + j] : j < f.n ? -x[(int64_t)faces[f.faces + j / b] * b + j % b] : 0.0;
    __syncthreads();
    const int lane = threadIdx.x % LPR, grp = threadIdx.x / LPR;
    for (int r = it.y + grp; r < min(it.y + R, f.nfs); r += G) {
        const double s = row_dot<LPR, TF>(fac + f.zv + (int64_t)r * ld, u, ld, lane);
        if (lane == 0) x[(int64_t)faces[f.faces + r / b] * b + r % b] = s;
    }
}

// lanes per row for rows of ncols values
inline int lanes_per_row(int ncols) { return ncols <= 64 ? 8 : ncols <= 160 ? 16 : ncols <= 640 ? 32 : 64; }

// ---------------------------------------------------------------------------
// host: symbolic phase
// ---------------------------------------------------------------------------
struct Symbolic {
    int F = 0, N = 0, b = 0;
    std::vector<Front> fronts;
    std::vector<Level> levels;           // by depth, 0 = root
    std::vector<int32_t> face_list, fmap, node_list;
    std::vector<int4> rec;
    std::vector<int2> fw_items, bw_items, gi_items;
    std::vector<int2> gsrc;              // per front row: the two children's contribution indices, or -1
    std::vector<int32_t> bsr_ptr, bsr_col;
    int64_t fac_size = 0, piv_size = 0, cvec_size = 0;
    int64_t flops = 0;
    int max_depth = 0;
};

template <typename T> T* dmalloc(size_t n)
{
    void* p = nullptr;
    TL_TRY(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
    return static_cast<T*>(p);
}

void build_symbolic(const pbte_trace_desc* d, Symbolic& S)
{
    const int F = d->n_faces, N = d->n_cells, b = d->blk;
    const int leaf = d->leaf_cells > 0 ? d->leaf_cells : 16;
    if (F <= 0 || N <= 0 || b <= 0) throw std::string("trace solve: empty system");
    S.F = F;
    S.N = N;
    S.b = b;
    S.bsr_ptr.assign(d->bsr_ptr, d->bsr_ptr + F + 1);
    S.bsr_col.assign(d->bsr_col, d->bsr_col + d->bsr_ptr[F]);
    // tree by recursive coordinate bisection; cells permuted so that every
    // node holds a contiguous range
    std::vector<int32_t> cell(N);
    std::iota(cell.begin(), cell.end(), 0);
    struct Node { int parent, depth, lo, hi, child[2]; };
    std::vector<Node> nodes;
    nodes.push_back({-1, 0, 0, N, {-1, -1}});
    for (size_t v = 0; v < nodes.size(); ++v) {
        const int lo = nodes[v].lo, hi = nodes[v].hi;
        if (hi - lo <= leaf) continue;
        double mn[2] = {1e300, 1e300}, mx[2] = {-1e300, -1e300};
        for (int k = lo; k < hi; ++k)
            for (int a = 0; a < 2; ++a) {
                mn[a] = std::min(mn[a], d->cell_xy[2 * cell[k] + a]);
                mx[a] = std::max(mx[a], d->cell_xy[2 * cell[k] + a]);
            }
        const int ax = (mx[1] - mn[1] > mx[0] - mn[0]) ? 1 : 0;
        const int mid = lo + (hi - lo) / 2;
        auto key = [&](int c) { return d->cell_xy[2 * c + ax]; };
        std::nth_element(cell.begin() + lo, cell.begin() + mid, cell.begin() + hi, [&](int p, int q) {
            const double kp = key(p), kq = key(q);
            return kp < kq || (kp == kq && p < q);
        });
        const int depth = nodes[v].depth + 1;
        const int a = (int)nodes.size();
        nodes.push_back({(int)v, depth, lo, mid, {-1, -1}});
        nodes.push_back({(int)v, depth, mid, hi, {-1, -1}});
        nodes[v].child[0] = a;
        nodes[v].child[1] = a + 1;
    }
    const int V = (int)nodes.size();
    std::vector<int32_t> leaf_of(N, -1);
    for (int v = 0; v < V; ++v)
        if (nodes[v].child[0] < 0)
            for (int k = nodes[v].lo; k < nodes[v].hi; ++k) leaf_of[cell[k]] = v;
    // owner of each face: the lowest node holding both cells it couples
    std::vector<int32_t> owner(F);
    for (int f = 0; f < F; ++f) {
        const int c1 = d->face_t1[f], c2 = d->face_t2[f];
        if (c1 < 0 || c1 >= N || c2 >= N) throw std::string("trace solve: face cell out of range");
        int a = leaf_of[c1], q = c2 >= 0 ? leaf_of[c2] : a;
        while (a != q) {
            if (nodes[a].depth >= nodes[q].depth) a = nodes[a].parent;
            else q = nodes[q].parent;
        }
        owner[f] = a;
    }
    // faces touching each cell: its own three and every face that couples it
    std::vector<int32_t> tptr(N + 1, 0), tf;
    {
        std::vector<std::vector<int32_t>> t(N);
        for (int c = 0; c < N; ++c)
            for (int k = 0; k < 3; ++k) t[c].push_back(d->tri_faces[3 * c + k]);
        for (int f = 0; f < F; ++f) {
            t[d->face_t1[f]].push_back(f);
            if (d->face_t2[f] >= 0) t[d->face_t2[f]].push_back(f);
        }
        for (int c = 0; c < N; ++c) {
            std::sort(t[c].begin(), t[c].end());
            t[c].erase(std::unique(t[c].begin(), t[c].end()), t[c].end());
            tptr[c + 1] = tptr[c] + (int)t[c].size();
            tf.insert(tf.end(), t[c].begin(), t[c].end());
        }
    }
    // own faces of each node (ascending face id)
    std::vector<std::vector<int32_t>> own(V), anc(V);
    for (int f = 0; f < F; ++f) own[owner[f]].push_back(f);
    std::vector<int32_t> stamp(F, -1);
    for (int v = 0; v < V; ++v) {
        const int dv = nodes[v].depth;
        for (int k = nodes[v].lo; k < nodes[v].hi; ++k) {
            const int c = cell[k];
            for (int e = tptr[c]; e < tptr[c + 1]; ++e) {
                const int f = tf[e];
                if (nodes[owner[f]].depth < dv && stamp[f] != v) {
                    stamp[f] = v;
                    anc[v].push_back(f);
                }
            }
        }
        std::sort(anc[v].begin(), anc[v].end());
    }
    // fronts, levels, offsets
    int max_depth = 0;
    for (auto& nd : nodes) max_depth = std::max(max_depth, nd.depth);
    S.max_depth = max_depth;
    S.levels.assign(max_depth + 1, Level());
    S.fronts.resize(V);
    int64_t fac = 0, piv = 0, cvec = 0;
    for (int v = 0; v < V; ++v) {
        Front& f = S.fronts[v];
        f.nfs = (int)own[v].size() * b;
        f.ncb = (int)anc[v].size() * b;
        f.n = f.nfs + f.ncb;
        f.faces = (int32_t)S.face_list.size();
        S.face_list.insert(S.face_list.end(), own[v].begin(), own[v].end());
        S.face_list.insert(S.face_list.end(), anc[v].begin(), anc[v].end());
        f.child[0] = nodes[v].child[0];
        f.child[1] = nodes[v].child[1];
        f.cmap[0] = f.cmap[1] = -1;
        f.zv = fac;
        fac += (int64_t)f.nfs * row_ld(f.n);
        f.w = fac;
        fac += (int64_t)f.ncb * row_ld(f.nfs);
        f.piv = (int32_t)piv;
        piv += f.nfs;
        f.cvec = (int32_t)cvec;
        cvec += f.ncb;
        Level& L = S.levels[nodes[v].depth];
        f.ws = L.ws_size;
        L.ws_size += (int64_t)f.n * f.n;
        f.pan = L.pan_size;
        L.pan_size += 2 * (int64_t)f.n * KB + KB * KB;
        L.nodes.push_back(v);
        L.n_max = std::max(L.n_max, f.n);
        L.nfs_max = std::max(L.nfs_max, f.nfs);
        L.ncb_max = std::max(L.ncb_max, f.ncb);
        S.flops += 2 * (int64_t)f.nfs * f.n * f.n;
    }
    if (piv != (int64_t)F * b) throw std::string("trace solve: faces not owned exactly once");
    S.fac_size = fac;
    S.piv_size = piv;
    S.cvec_size = cvec;
    // child maps: positions of the child's ancestor faces in the parent front
    std::vector<int32_t> pos(F, -1);
    for (int v = 0; v < V; ++v) {
        Front& f = S.fronts[v];
        const int nfc = f.n / b;
        for (int k = 0; k < nfc; ++k) pos[S.face_list[f.faces + k]] = k;
        f.gsrc = (int32_t)S.gsrc.size();
        S.gsrc.resize(S.gsrc.size() + f.n, make_int2(-1, -1));
        for (int s = 0; s < 2; ++s) {
            const int c = f.child[s];
            if (c < 0) continue;
            const Front& ch = S.fronts[c];
            f.cmap[s] = (int32_t)S.fmap.size();
            const int nown = ch.nfs / b;
            for (int k = 0; k < ch.ncb / b; ++k) {
                const int g = S.face_list[ch.faces + nown + k];
                if (pos[g] < 0) throw std::string("trace solve: child face missing from parent front");
                S.fmap.push_back(pos[g]);
                for (int e = 0; e < b; ++e) {
                    int2& src = S.gsrc[(size_t)f.gsrc + (size_t)pos[g] * b + e];
                    (s == 0 ? src.x : src.y) = ch.cvec + k * b + e;
                }
            }
            S.levels[nodes[v].depth].has_children = true;
            S.levels[nodes[v].depth].ncb_child_max = std::max(S.levels[nodes[v].depth].ncb_child_max, ch.ncb);
        }
        for (int k = 0; k < nfc; ++k) pos[S.face_list[f.faces + k]] = -1;
    }
    // original blocks, grouped by destination node and level
    std::vector<std::vector<int4>> recs(max_depth + 1);
    {
        std::vector<std::vector<std::pair<int, int>>> by_node(V);   // (block, row face)
        for (int f = 0; f < F; ++f)
            for (int e = d->bsr_ptr[f]; e < d->bsr_ptr[f + 1]; ++e) {
                const int g = d->bsr_col[e];
                if (g < 0 || g >= F) throw std::string("trace solve: block column out of range");
                const int of = owner[f], og = owner[g];
                const int dst = nodes[of].depth >= nodes[og].depth ? of : og;
                by_node[dst].push_back({e, f});
            }
        for (int v = 0; v < V; ++v) {
            const Front& fv = S.fronts[v];
            const int nfc = fv.n / b;
            for (int k = 0; k < nfc; ++k) pos[S.face_list[fv.faces + k]] = k;
            for (auto [e, f] : by_node[v]) {
                const int g = d->bsr_col[e];
                if (pos[f] < 0 || pos[g] < 0)
                    throw std::string("trace solve: a block couples faces that share no cell");
                recs[nodes[v].depth].push_back({e, v, pos[f] * b, pos[g] * b});
            }
            for (int k = 0; k < nfc; ++k) pos[S.face_list[fv.faces + k]] = -1;
        }
    }
    for (int l = 0; l <= max_depth; ++l) {
        Level& L = S.levels[l];
        L.node_off = (int32_t)S.node_list.size();
        S.node_list.insert(S.node_list.end(), L.nodes.begin(), L.nodes.end());
        L.rec_off = (int32_t)S.rec.size();
        L.rec_n = (int32_t)recs[l].size();
        S.rec.insert(S.rec.end(), recs[l].begin(), recs[l].end());
        L.fw_off = (int32_t)S.fw_items.size();
        L.bw_off = (int32_t)S.bw_items.size();
        L.gi_off = (int32_t)S.gi_items.size();
        for (int v : L.nodes)
            for (int r = 0; r < S.fronts[v].n; r += 256) S.gi_items.push_back({v, r});
        L.gi_n = (int32_t)S.gi_items.size() - L.gi_off;
        int nfs_max = 0, n_max = 0;
        for (int v : L.nodes) {
            nfs_max = std::max(nfs_max, S.fronts[v].nfs);
            n_max = std::max(n_max, S.fronts[v].n);
        }
        L.lpr_fw = lanes_per_row(nfs_max);
        L.lpr_bw = lanes_per_row(n_max);
        const int rows_fw = 2 * 256 / L.lpr_fw, rows_bw = 2 * 256 / L.lpr_bw;
        for (int v : L.nodes) {
            const Front& f = S.fronts[v];
            for (int r = 0; r < f.ncb; r += rows_fw) S.fw_items.push_back({v, r});
            for (int r = 0; r < f.nfs; r += rows_bw) S.bw_items.push_back({v, r});
        }
        L.fw_n = (int32_t)S.fw_items.size() - L.fw_off;
        L.bw_n = (int32_t)S.bw_items.size() - L.bw_off;
        if ((int64_t)row_ld(L.n_max) * 8 > 64 * 1024)
            throw std::string("trace solve: a front exceeds the LDS vector of the solve (n > 8192)");
    }
}

}  // namespace

struct pbte_trace {
    Symbolic S;
    hipStream_t stream = nullptr;
    std::string err;
    bool factored = false;
    int pivoting = 0;         // 0 diagonal-block pivots, 1 partial pivoting
    int force_pivoting = -1;  // -1 automatic, else the mode to use
    double berr = 0.0;        // backward error of the factors on the test vector
    // device
    Front* d_fr = nullptr;
    int32_t *d_faces = nullptr, *d_fmap = nullptr, *d_nodes = nullptr, *d_perm = nullptr, *d_ipiv = nullptr,
            *d_status = nullptr;
    int4* d_rec = nullptr;
    int2 *d_fw = nullptr, *d_bw = nullptr, *d_gi = nullptr, *d_gsrc = nullptr;
    double *d_fac = nullptr, *d_ws[2] = {nullptr, nullptr}, *d_pan = nullptr, *d_y = nullptr, *d_cv = nullptr;
    float* d_fac32 = nullptr;   // binary32 copy of the factors (made by the first binary32 solve)
    bool fac32_current = false;
    int64_t ws_size[2] = {0, 0}, pan_size = 0;
    // the solve's launches (two per level forward, one backward) recorded
    // once per (precision, rhs, x) as a HIP graph on a private stream and
    // replayed on `stream`; dropped when the factors change
    struct SolveGraph {
        int precision;
        const double* rhs;
        double* x;
        hipGraphExec_t exec;
    };
    std::vector<SolveGraph> graphs;
    hipStream_t cap_stream = nullptr;
    bool use_graphs = true;
    static constexpr size_t kMaxGraphs = 8;

    void drop_graphs()
    {
        for (auto& g : graphs) (void)hipGraphExecDestroy(g.exec);
        graphs.clear();
    }

    ~pbte_trace()
    {
        drop_graphs();
        if (cap_stream) (void)hipStreamDestroy(cap_stream);
        void* ps[] = {d_fr, d_faces, d_fmap, d_nodes, d_perm, d_ipiv, d_status, d_rec, d_fw, d_bw, d_gi, d_gsrc,
                      d_fac, d_ws[0], d_ws[1], d_pan, d_y, d_cv, d_fac32};
        for (void* p : ps)
            if (p) (void)hipFree(p);
    }

    void alloc()
    {
        if (d_fr) return;
        for (size_t l = 0; l < S.levels.size(); ++l) {
            ws_size[l & 1] = std::max(ws_size[l & 1], S.levels[l].ws_size);
            pan_size = std::max(pan_size, S.levels[l].pan_size);
        }
        auto up = [&](auto*& dst, const auto& v) {
            using E = std::remove_reference_t<decltype(v[0])>;
            dst = dmalloc<std::remove_const_t<E>>(v.size());
            if (!v.empty())
                TL_TRY(hipMemcpyAsync(dst, v.data(), v.size() * sizeof(E), hipMemcpyHostToDevice, stream));
        };
        up(d_fr, S.fronts);
        up(d_faces, S.face_list);
        up(d_fmap, S.fmap);
        up(d_nodes, S.node_list);
        up(d_rec, S.rec);
        up(d_fw, S.fw_items);
        up(d_bw, S.bw_items);
        up(d_gi, S.gi_items);
        up(d_gsrc, S.gsrc);
        d_perm = dmalloc<int32_t>(S.piv_size);
        d_ipiv = dmalloc<int32_t>(S.piv_size);
        d_status = dmalloc<int32_t>(1);
        d_fac = dmalloc<double>(S.fac_size);
        d_ws[0] = dmalloc<double>(ws_size[0]);
        d_ws[1] = dmalloc<double>(ws_size[1]);
        d_pan = dmalloc<double>(pan_size);
        // y is read past a front's own values up to its padded row length
        // (against the zero padding of W): four zeros after the last one
        d_y = dmalloc<double>(S.piv_size + 4);
        hipLaunchKernelGGL(k_fill, dim3(64), dim3(256), 0, stream, d_y, S.piv_size + 4);
        TL_TRY(hipGetLastError());
        d_cv = dmalloc<double>(S.cvec_size);
        TL_TRY(hipStreamSynchronize(stream));
    }
};

namespace {
inline unsigned cdiv(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

template <typename Fn>
int tguard(pbte_trace* t, Fn&& fn)
{
    if (!t) return -1;
    try {
        return fn();
    } catch (const std::string& e) {
        t->err = e;
        return -2;
    }
}
}  // namespace

extern "C" pbte_trace* pbte_trace_create(const pbte_trace_desc* desc, void* stream, char* err, int32_t errlen)
{
    auto* t = new pbte_trace();
    try {
        if (!desc) throw std::string("null descriptor");
        t->stream = static_cast<hipStream_t>(stream);
        build_symbolic(desc, t->S);
        return t;
    } catch (const std::string& e) {
        if (err && errlen > 0) {
            std::strncpy(err, e.c_str(), (size_t)errlen - 1);
            err[errlen - 1] = 0;
        }
        delete t;
        return nullptr;
    }
}

extern "C" void pbte_trace_destroy(pbte_trace* t) { delete t; }

extern "C" const char* pbte_trace_last_error(pbte_trace* t) { return t ? t->err.c_str() : "null trace solver"; }

extern "C" int pbte_trace_graphs(pbte_trace* t, int32_t on)
{
    return tguard(t, [&] {
        t->use_graphs = on != 0;
        if (!t->use_graphs) t->drop_graphs();
        return 0;
    });
}

extern "C" int pbte_trace_options(pbte_trace* t, int32_t pivoting)
{
    return tguard(t, [&] {
        if (pivoting < -1 || pivoting > 1) throw std::string("trace solve: pivoting must be -1, 0 or 1");
        t->force_pivoting = pivoting;
        return 0;
    });
}

extern "C" int pbte_trace_quality(pbte_trace* t, int32_t* pivoting, double* berr)
{
    return tguard(t, [&] {
        if (!t->factored) throw std::string("trace solve: not factorised");
        *pivoting = t->pivoting;
        *berr = t->berr;
        return 0;
    });
}

extern "C" int pbte_trace_level(pbte_trace* t, int32_t level, int64_t* out8)
{
    return tguard(t, [&] {
        const Symbolic& S = t->S;
        if (level < 0 || level >= (int)S.levels.size()) throw std::string("trace level out of range");
        const Level& L = S.levels[level];
        int64_t nfs = 0, ncb = 0, zv = 0, w = 0;
        for (int v : L.nodes) {
            const Front& f = S.fronts[v];
            nfs += f.nfs;
            ncb += f.ncb;
            zv += (int64_t)f.nfs * row_ld(f.n);
            w += (int64_t)f.ncb * row_ld(f.nfs);
        }
        out8[0] = (int64_t)L.nodes.size();
        out8[1] = L.n_max;
        out8[2] = nfs;
        out8[3] = ncb;
        out8[4] = zv;
        out8[5] = w;
        out8[6] = L.fw_n;
        out8[7] = L.bw_n;
        return 0;
    });
}

extern "C" int pbte_trace_info(pbte_trace* t, int64_t* out8)
{
    return tguard(t, [&] {
        const Symbolic& S = t->S;
        int64_t ws = 0, nmax = 0;
        for (size_t l = 0; l < S.levels.size(); ++l) nmax = std::max<int64_t>(nmax, S.levels[l].n_max);
        int64_t wsz[2] = {0, 0}, pz = 0;
        for (size_t l = 0; l < S.levels.size(); ++l) {
            wsz[l & 1] = std::max(wsz[l & 1], S.levels[l].ws_size);
            pz = std::max(pz, S.levels[l].pan_size);
        }
        ws = wsz[0] + wsz[1] + pz;
        out8[0] = (int64_t)S.fronts.size();
        out8[1] = (int64_t)S.levels.size();
        out8[2] = nmax;
        out8[3] = S.fac_size;
        out8[4] = S.flops;
        out8[5] = ws * 8;
        out8[6] = (int64_t)S.F * S.b;
        out8[7] = (int64_t)S.rec.size();
        return 0;
    });
}

namespace {
// numeric phase with diagonal-block pivoting (block_pivots) or with partial
// pivoting among each front's fully summed rows; returns the status word
int factor_pass(pbte_trace* t, const double* bsr_val, bool block_pivots)
{
    Symbolic& S = t->S;
    hipStream_t st = t->stream;
    const int big = 0x7fffffff;
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, st, t->d_status, big);
    for (int l = S.max_depth; l >= 0; --l) {
        const Level& L = S.levels[l];
        const int nf = (int)L.nodes.size();
        const int32_t* nodes = t->d_nodes + L.node_off;
        double* ws = t->d_ws[l & 1];
        double* wsc = t->d_ws[(l + 1) & 1];
        hipLaunchKernelGGL(k_fill, dim3(std::min<int64_t>(cdiv(L.ws_size, 256), 65536)), dim3(256), 0, st, ws,
                           L.ws_size);
        if (L.rec_n > 0) {
            const int64_t tot = (int64_t)L.rec_n * S.b * S.b;
            hipLaunchKernelGGL(k_assemble, dim3(std::min<int64_t>(cdiv(tot, 256), 65536)), dim3(256), 0, st,
                               t->d_rec + L.rec_off, L.rec_n, bsr_val, S.b, t->d_fr, ws);
        }
        if (L.has_children)
            for (int s = 0; s < 2; ++s)
                hipLaunchKernelGGL(k_extend, dim3(nf, cdiv(L.ncb_child_max, 8)), dim3(256), 0, st, nodes, s,
                                   t->d_fr, t->d_fmap, S.b, ws, wsc);
        const int tpr = cdiv(L.n_max, TILE);
        for (int j0 = 0; j0 < L.nfs_max; j0 += KB) {
            if (block_pivots) {
                hipLaunchKernelGGL(k_block_inv, dim3(nf), dim3(256), 0, st, nodes, t->d_fr, ws, t->d_pan, j0,
                                   t->d_perm, t->d_status);
                hipLaunchKernelGGL(k_panel_apply, dim3(nf, cdiv(L.n_max, 64)), dim3(256), 0, st, nodes, t->d_fr, ws,
                                   t->d_pan, j0);
                hipLaunchKernelGGL(k_copy_rows, dim3(nf, cdiv(L.n_max, 256)), dim3(256), 0, st, nodes, t->d_fr, ws,
                                   t->d_pan, j0);
            } else {
                hipLaunchKernelGGL(k_panel, dim3(nf), dim3(1024), 0, st, nodes, t->d_fr, ws, t->d_pan, j0,
                                   t->d_ipiv, t->d_perm, t->d_status);
                hipLaunchKernelGGL(k_swap_rows, dim3(nf, cdiv(L.n_max, 256)), dim3(256), 0, st, nodes, t->d_fr, ws,
                                   t->d_pan, j0, t->d_ipiv);
            }
            hipLaunchKernelGGL(k_update, dim3(tpr * tpr, nf), dim3(256), 0, st, nodes, t->d_fr, ws, t->d_pan, j0,
                               tpr);
        }
        hipLaunchKernelGGL(k_extract, dim3(nf, cdiv(L.n_max, 4)), dim3(256), 0, st, nodes, t->d_fr, ws, t->d_fac);
    }
    TL_TRY(hipGetLastError());
    int status = big;
    TL_TRY(hipMemcpyAsync(&status, t->d_status, sizeof(int), hipMemcpyDeviceToHost, st));
    TL_TRY(hipStreamSynchronize(st));
    return status == big ? 0 : 1;
}

void solve_pass(pbte_trace* t, const double* rhs, double* x);

// normwise backward error max|K x - b| / (max_i sum_j |K_ij| max|x| + max|b|)
// of the factors on a fixed test vector
double backward_error(pbte_trace* t, const double* bsr_val)
{
    Symbolic& S = t->S;
    hipStream_t st = t->stream;
    const int64_t n = (int64_t)S.F * S.b;
    std::vector<double> hb(n), hx(n), hr(n), ha(n);
    uint64_t z = 0x9e3779b97f4a7c15ull;
    for (int64_t i = 0; i < n; ++i) {     // splitmix64, uniform in [-1, 1)
        z += 0x9e3779b97f4a7c15ull;
        uint64_t q = z;
        q = (q ^ (q >> 30)) * 0xbf58476d1ce4e5b9ull;
        q = (q ^ (q >> 27)) * 0x94d049bb133111ebull;
        q ^= q >> 31;
        hb[i] = (double)(q >> 11) * 0x1.0p-52 - 1.0;
    }
    double *db = dmalloc<double>(n), *dx = dmalloc<double>(n), *dr = dmalloc<double>(n), *da = dmalloc<double>(n);
    int32_t *dptr = dmalloc<int32_t>(S.F + 1), *dcol = dmalloc<int32_t>(S.bsr_col.size());
    auto cleanup = [&] {
        for (void* p : {(void*)db, (void*)dx, (void*)dr, (void*)da, (void*)dptr, (void*)dcol}) (void)hipFree(p);
    };
    try {
        TL_TRY(hipMemcpyAsync(db, hb.data(), n * 8, hipMemcpyHostToDevice, st));
        TL_TRY(hipMemcpyAsync(dptr, S.bsr_ptr.data(), (S.F + 1) * 4, hipMemcpyHostToDevice, st));
        TL_TRY(hipMemcpyAsync(dcol, S.bsr_col.data(), S.bsr_col.size() * 4, hipMemcpyHostToDevice, st));
        solve_pass(t, db, dx);
        hipLaunchKernelGGL(k_bsr_residual, dim3(cdiv(n, 256)), dim3(256), 0, st, dptr, dcol, bsr_val, S.F, S.b, dx,
                           db, dr, da);
        TL_TRY(hipGetLastError());
        TL_TRY(hipMemcpyAsync(hx.data(), dx, n * 8, hipMemcpyDeviceToHost, st));
        TL_TRY(hipMemcpyAsync(hr.data(), dr, n * 8, hipMemcpyDeviceToHost, st));
        TL_TRY(hipMemcpyAsync(ha.data(), da, n * 8, hipMemcpyDeviceToHost, st));
        TL_TRY(hipStreamSynchronize(st));
    } catch (...) {
        cleanup();
        throw;
    }
    cleanup();
    double rmax = 0, xmax = 0, amax = 0, bmax = 0;
    bool finite = true;
    for (int64_t i = 0; i < n; ++i) {
        finite = finite && std::isfinite(hx[i]) && std::isfinite(hr[i]);
        rmax = std::max(rmax, std::fabs(hr[i]));
        xmax = std::max(xmax, std::fabs(hx[i]));
        amax = std::max(amax, ha[i]);
        bmax = std::max(bmax, std::fabs(hb[i]));
    }
    return finite ? rmax / (amax * xmax + bmax) : INFINITY;
}
}  // namespace

extern "C" int pbte_trace_factor(pbte_trace* t, const double* bsr_val)
{
    return tguard(t, [&] {
        t->alloc();
        t->drop_graphs();
        t->factored = false;
        t->fac32_current = false;
        // diagonal-block pivots first; partial pivoting among the fully
        // summed rows when a block has no pivot or the factors fail the
        // backward-error test (or as the caller forces)
        if (t->force_pivoting == 0) {
            if (factor_pass(t, bsr_val, true) != 0) {
                t->err = "trace solve: a diagonal block without a nonzero pivot";
                return 1;
            }
            t->factored = true;
            t->berr = backward_error(t, bsr_val);
            t->pivoting = 0;
            return 0;
        }
        if (t->force_pivoting < 0 && factor_pass(t, bsr_val, true) == 0) {
            t->factored = true;
            t->berr = backward_error(t, bsr_val);
            if (t->berr <= 1e-13) {
                t->pivoting = 0;
                return 0;
            }
        }
        t->factored = false;
        if (factor_pass(t, bsr_val, false) != 0) {
            t->err = "trace solve: the trace matrix is singular (a front without a nonzero pivot)";
            return 1;
        }
        t->factored = true;
        t->berr = backward_error(t, bsr_val);
        t->pivoting = 1;
        return 0;
    });
}

namespace {
// lanes per row of a solve product: the level's choice for binary64 rows;
// binary32 rows hold half the bytes, so half the lanes give each lane the
// same 16-byte loads per row (the work items, 2 * 256 / lpr rows, stay)
template <typename TF> int solve_lanes(int lpr) { return sizeof(TF) == 4 ? std::max(4, lpr / 2) : lpr; }

template <typename TF>
void solve_pass_t(pbte_trace* t, const TF* fac, const double* rhs, double* x, hipStream_t st)
{
    Symbolic& S = t->S;
    for (int l = S.max_depth; l >= 0; --l) {
        const Level& L = S.levels[l];
        hipLaunchKernelGGL(k_fwd_gather, dim3(L.gi_n), dim3(256), 0, st, t->d_gi + L.gi_off, t->d_fr, t->d_faces,
                           t->d_gsrc, t->d_perm, S.b, rhs, t->d_y, t->d_cv);
        if (L.fw_n > 0) {
            const int lpr = solve_lanes<TF>(L.lpr_fw);
            auto fn = lpr == 4 ? k_fwd_gemv<4, TF> : lpr == 8 ? k_fwd_gemv<8, TF> : lpr == 16 ? k_fwd_gemv<16, TF>
                    : lpr == 32 ? k_fwd_gemv<32, TF> : k_fwd_gemv<64, TF>;
            hipLaunchKernelGGL(fn, dim3(L.fw_n), dim3(256), 0, st, t->d_fw + L.fw_off, t->d_fr, fac, t->d_y,
                               t->d_cv, 2 * 256 / L.lpr_fw);
        }
    }
    for (int l = 0; l <= S.max_depth; ++l) {
        const Level& L = S.levels[l];
        if (L.bw_n > 0) {
            const int lpr = solve_lanes<TF>(L.lpr_bw);
            auto fn = lpr == 4 ? k_bwd<4, TF> : lpr == 8 ? k_bwd<8, TF> : lpr == 16 ? k_bwd<16, TF>
                    : lpr == 32 ? k_bwd<32, TF> : k_bwd<64, TF>;
            hipLaunchKernelGGL(fn, dim3(L.bw_n), dim3(256), (size_t)row_ld(L.n_max) * 8, st, t->d_bw + L.bw_off, t->d_fr,
                               t->d_faces, S.b, fac, t->d_y, x, 2 * 256 / L.lpr_bw);
        }
    }
    TL_TRY(hipGetLastError());
}

void solve_pass(pbte_trace* t, const double* rhs, double* x) { solve_pass_t<double>(t, t->d_fac, rhs, x, t->stream); }

// one solve on t->stream: replayed from its graph when one was recorded
// for these buffers, else recorded now (at most kMaxGraphs per factors;
// past that, or with graphs off, launched directly).  Recording captures
// the launches on a private stream without running them; the graph then
// runs on t->stream in its order like any launch.
template <typename TF>
void solve_graphed(pbte_trace* t, int precision, const TF* fac, const double* rhs, double* x)
{
    if (t->use_graphs) {
        for (const auto& g : t->graphs)
            if (g.precision == precision && g.rhs == rhs && g.x == x) {
                TL_TRY(hipGraphLaunch(g.exec, t->stream));
                return;
            }
        if (t->graphs.size() < pbte_trace::kMaxGraphs) {
            if (!t->cap_stream) TL_TRY(hipStreamCreateWithFlags(&t->cap_stream, hipStreamNonBlocking));
            hipGraph_t graph = nullptr;
            TL_TRY(hipStreamBeginCapture(t->cap_stream, hipStreamCaptureModeRelaxed));
            solve_pass_t<TF>(t, fac, rhs, x, t->cap_stream);
            TL_TRY(hipStreamEndCapture(t->cap_stream, &graph));
            hipGraphExec_t exec = nullptr;
            const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            TL_TRY(e);
            t->graphs.push_back({precision, rhs, x, exec});
            TL_TRY(hipGraphLaunch(exec, t->stream));
            return;
        }
    }
    solve_pass_t<TF>(t, fac, rhs, x, t->stream);
}
}  // namespace

extern "C" int pbte_trace_solve(pbte_trace* t, int32_t precision, const double* rhs, double* x)
{
    return tguard(t, [&] {
        if (!t->factored) throw std::string("trace solve: pbte_trace_factor has not succeeded");
        if (precision == PBTE_FP64) {
            solve_graphed<double>(t, PBTE_FP64, t->d_fac, rhs, x);
        } else if (precision == PBTE_FP32) {
            if (!t->d_fac32) t->d_fac32 = dmalloc<float>((size_t)t->S.fac_size);
            if (!t->fac32_current) {
                hipLaunchKernelGGL(k_narrow, dim3(4096), dim3(256), 0, t->stream, t->d_fac, t->d_fac32,
                                   t->S.fac_size);
                TL_TRY(hipGetLastError());
                t->fac32_current = true;
            }
            solve_graphed<float>(t, PBTE_FP32, t->d_fac32, rhs, x);
        } else {
            throw std::string("trace solve: precision is PBTE_FP64 or PBTE_FP32");
        }
        return 0;
    });
}
