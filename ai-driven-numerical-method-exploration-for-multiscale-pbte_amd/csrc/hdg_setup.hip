// HDG synthetic-system operators of GSIS, built on the device.
//
// Replaces the host setup of the reference's acceleration package
// (solver-python/pybte/acceleration/assembly.py:44-268):
//
// * build_local_operators: per element the 7P x 7P matrix AA_SOL of the
//   seven macroscopic fields [T, qx, qy, Lxx, Lxy, Lyx, Lyy] (blocks of the
//   element integral tensors), its inverse, and AA_TRACE = inv(AA_SOL) times
//   the 7P x 3Q face coupling of each local face;
// * build_ba_sol: per (element, local face) the 3Q x 7P trace rows
//   (average plus stabilised jump on interior and periodic faces, the
//   momentum rows of a thermalising wall for variant A, zero otherwise);
// * assemble_global_matrix: the trace matrix K, block row per face: the
//   int_fc_fc block of each field on the diagonal minus BA_SOL AA_TRACE
//   from each side, rows mirrored for the partner side of a reversed
//   periodic pair, summed in the order diagonal, side 1 (local faces 0..2),
//   side 2, and entries with |K| <= drop_tol set to zero.
//
// One workgroup per element builds and inverts AA_SOL in LDS (Gauss-Jordan
// with row interchanges, largest magnitude pivot, first on ties) and forms
// AA_TRACE; one thread per entry of a K block forms its sum.  All binary64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "pbte_hip.h"

namespace {

struct HdgArgs {
    int N, P, Q, F, variant_b;
    double cv, tau_r, tc5, scale, st0, st1, st2, drop_tol;
    const double *tt, *tx, *ty, *tfc, *tfcfc, *fcfc, *nrm;
    const int32_t* kind;
};

HdgArgs make_args(const pbte_hdg_desc* d)
{
    HdgArgs a;
    a.N = d->n_tris;
    a.P = d->nd;
    a.Q = d->nf;
    a.F = d->n_faces;
    a.variant_b = d->variant_b;
    a.cv = d->cv;
    a.tau_r = d->tau_r;
    a.scale = d->variant_b ? d->tau_r : 1.0;
    a.tc5 = d->tau_c * a.scale / 5.0;
    a.st0 = d->st[0];
    a.st1 = d->st[1];
    a.st2 = d->st[2];
    a.drop_tol = d->drop_tol;
    a.tt = d->int_tri_tri;
    a.tx = d->int_tri_tri_x;
    a.ty = d->int_tri_tri_y;
    a.tfc = d->int_tri_fc;
    a.tfcfc = d->int_tri_fc_fc;
    a.fcfc = d->int_fc_fc;
    a.nrm = d->tri_normal;
    a.kind = d->face_kind;
    return a;
}

// entry (m, l) of block (R, C) of AA_SOL for element i (zero outside the
// blocks the system couples); the element tensors are aa = int_tri_tri_x,
// bb = int_tri_tri_y, cc = aa^T, dd = bb^T, pp = sum over faces of
// int_tri_fc, oo = int_tri_tri[., ., i]
__device__ double aa_sol_entry(const HdgArgs& a, int i, int R, int C, int m, int l)
{
    const int P = a.P;
    const size_t NPP = (size_t)i * P * P;
    auto aa = [&](int r, int c) { return a.tx[NPP + (size_t)r * P + c]; };
    auto bb = [&](int r, int c) { return a.ty[NPP + (size_t)r * P + c]; };
    auto oo = [&](int r, int c) { return a.tt[((size_t)r * P + c) * a.N + i]; };
    auto pp = [&](int r, int c) {
        const size_t b = (size_t)i * 3 * P * P + (size_t)r * P + c;
        return (a.tfc[b] + a.tfc[b + (size_t)P * P]) + a.tfc[b + 2 * (size_t)P * P];
    };
    const double cv3 = a.cv;
    switch (R * 7 + C) {
    case 0 * 7 + 0: return a.st0 * pp(m, l);
    case 0 * 7 + 1: return aa(l, m);
    case 0 * 7 + 2: return bb(l, m);
    case 1 * 7 + 0: return -aa(m, l) * cv3 / 3.0 * a.scale;
    case 1 * 7 + 1: return a.variant_b ? oo(m, l) + a.st1 * pp(m, l) : a.st1 * pp(m, l) + oo(m, l) / a.tau_r;
    case 1 * 7 + 3: return -4.0 / 3.0 * aa(l, m);
    case 1 * 7 + 4: return -bb(l, m);
    case 1 * 7 + 5: return -bb(l, m);
    case 1 * 7 + 6: return 2.0 / 3.0 * aa(l, m);
    case 2 * 7 + 0: return -bb(m, l) * cv3 / 3.0 * a.scale;
    case 2 * 7 + 2: return a.variant_b ? oo(m, l) + a.st2 * pp(m, l) : a.st2 * pp(m, l) + oo(m, l) / a.tau_r;
    case 2 * 7 + 3: return 2.0 / 3.0 * bb(l, m);
    case 2 * 7 + 4: return -aa(l, m);
    case 2 * 7 + 5: return -aa(l, m);
    case 2 * 7 + 6: return -4.0 / 3.0 * bb(l, m);
    case 3 * 7 + 1: return aa(m, l) * a.tc5;
    case 3 * 7 + 3: return oo(m, l);
    case 4 * 7 + 1: return bb(m, l) * a.tc5;
    case 4 * 7 + 4: return oo(m, l);
    case 5 * 7 + 2: return aa(m, l) * a.tc5;
    case 5 * 7 + 5: return oo(m, l);
    case 6 * 7 + 2: return bb(m, l) * a.tc5;
    case 6 * 7 + 6: return oo(m, l);
    default: return 0.0;
    }
}

// entry (p, q) of block (R, C) of the 7P x 3Q face coupling of local face
// lf; ww = int_tri_fc_fc[., ., i, lf]
__device__ double face_coupling_entry(const HdgArgs& a, int i, int lf, int R, int C, int p, int q)
{
    const double ww = a.tfcfc[(((size_t)p * a.Q + q) * a.N + i) * 3 + lf];
    const double nx = a.nrm[((size_t)i * 3 + lf) * 2], ny = a.nrm[((size_t)i * 3 + lf) * 2 + 1];
    switch (R * 3 + C) {
    case 0 * 3 + 0: return a.st0 * ww;
    case 1 * 3 + 0: return -nx * ww * a.cv / 3.0 * a.scale;
    case 1 * 3 + 1: return a.st1 * ww;
    case 2 * 3 + 0: return -ny * ww * a.cv / 3.0 * a.scale;
    case 2 * 3 + 2: return a.st2 * ww;
    case 3 * 3 + 1: return nx * ww * a.tc5;
    case 4 * 3 + 1: return ny * ww * a.tc5;
    case 5 * 3 + 2: return nx * ww * a.tc5;
    case 6 * 3 + 2: return ny * ww * a.tc5;
    default: return 0.0;
    }
}

// entry (q, p) of block (R, C) of BA_SOL for (element i, local face lf);
// bb2 = int_tri_fc_fc[p, q, i, lf]; kind 0 interior / periodic, 1
// thermalising wall, 2 other walls (rows zero)
__device__ double ba_sol_entry(const HdgArgs& a, int i, int lf, int R, int C, int q, int p)
{
    const int kind = a.kind[i * 3 + lf];
    if (kind == 2) return 0.0;
    const double bb2 = a.tfcfc[(((size_t)p * a.Q + q) * a.N + i) * 3 + lf];
    const double nx = a.nrm[((size_t)i * 3 + lf) * 2], ny = a.nrm[((size_t)i * 3 + lf) * 2 + 1];
    const double s0 = a.st0, s1 = a.st1, s2 = a.st2;
    if (kind == 1) {
        if (a.variant_b) return 0.0;
        switch (R * 7 + C) {
        case 1 * 7 + 1: return bb2;
        case 1 * 7 + 3: return -4.0 * nx / 3.0 / s1 * bb2;
        case 1 * 7 + 4: return -ny / s1 * bb2;
        case 1 * 7 + 5: return -ny / s1 * bb2;
        case 1 * 7 + 6: return 2.0 * nx / 3.0 / s1 * bb2;
        case 2 * 7 + 2: return bb2;
        case 2 * 7 + 3: return 2.0 * ny / 3.0 / s2 * bb2;
        case 2 * 7 + 4: return -nx / s2 * bb2;
        case 2 * 7 + 5: return -nx / s2 * bb2;
        case 2 * 7 + 6: return -4.0 * ny / 3.0 / s2 * bb2;
        default: return 0.0;
        }
    }
    switch (R * 7 + C) {
    case 0 * 7 + 0: return 0.5 * bb2;
    case 0 * 7 + 1: return nx / 2.0 / s0 * bb2;
    case 0 * 7 + 2: return ny / 2.0 / s0 * bb2;
    case 1 * 7 + 1: return 0.5 * bb2;
    case 1 * 7 + 3: return -2.0 * nx / 3.0 / s1 * bb2;
    case 1 * 7 + 4: return -ny / 2.0 / s1 * bb2;
    case 1 * 7 + 5: return -ny / 2.0 / s1 * bb2;
    case 1 * 7 + 6: return nx / 3.0 / s1 * bb2;
    case 2 * 7 + 2: return 0.5 * bb2;
    case 2 * 7 + 3: return ny / 3.0 / s2 * bb2;
    case 2 * 7 + 4: return -nx / 2.0 / s2 * bb2;
    case 2 * 7 + 5: return -nx / 2.0 / s2 * bb2;
    case 2 * 7 + 6: return -2.0 * ny / 3.0 / s2 * bb2;
    default: return 0.0;
    }
}

// One workgroup per element: AA_SOL in LDS, inverted in place by the
// Gauss-Jordan sweep with row interchanges, then inv (optional) and
// AA_TRACE = inv T_lf for the three local faces.  Dynamic LDS:
// (7P)^2 + 7P * 3Q doubles.
__global__ void __launch_bounds__(256) k_local_ops(HdgArgs a, double* __restrict__ inv_out,
                                                   double* __restrict__ aa_trace, int32_t* __restrict__ nsing)
{
    extern __shared__ double lds[];
    __shared__ double s_val[256];
    __shared__ int s_idx[256];
    __shared__ int s_perm[128];
    const int i = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
    const int P = a.P, n = 7 * P, nq = 3 * a.Q;
    double* A = lds;
    double* T = lds + n * n;
    for (int e = tid; e < n * n; e += nt) {
        const int r = e / n, c = e - (e / n) * n;
        A[e] = aa_sol_entry(a, i, r / P, c / P, r % P, c % P);
    }
    for (int k = tid; k < n; k += nt) s_perm[k] = k;
    __syncthreads();
    bool singular = false;
    for (int k = 0; k < n && !singular; ++k) {
        double best = -1.0;
        int bi = k;
        for (int r = k + tid; r < n; r += nt) {
            const double v = fabs(A[r * n + k]);
            if (v > best) { best = v; bi = r; }
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
        singular = !(s_val[0] > 0.0);
        __syncthreads();
        if (singular) break;
        if (p != k) {
            for (int c = tid; c < n; c += nt) {
                const double t = A[k * n + c];
                A[k * n + c] = A[p * n + c];
                A[p * n + c] = t;
            }
            if (tid == 0) {
                const int t = s_perm[k];
                s_perm[k] = s_perm[p];
                s_perm[p] = t;
            }
        }
        __syncthreads();
        const double inv = 1.0 / A[k * n + k];
        for (int e = tid; e < n * n; e += nt) {
            const int r = e / n, c = e - (e / n) * n;
            if (r == k || c == k) continue;
            A[e] = fma(-(A[r * n + k] * inv), A[k * n + c], A[e]);
        }
        __syncthreads();
        for (int c = tid; c < n; c += nt)
            if (c != k) A[k * n + c] *= inv;
        for (int r = tid; r < n; r += nt)
            if (r != k) A[r * n + k] = -A[r * n + k] * inv;
        if (tid == 0) A[k * n + k] = inv;
        __syncthreads();
    }
    if (singular) {
        if (tid == 0) atomicAdd(nsing, 1);
        return;
    }
    // the sweep of the row-permuted matrix left inv(A) with its columns
    // permuted: inv(A)[r][perm[c]] = A[r][c]
    if (inv_out)
        for (int e = tid; e < n * n; e += nt) {
            const int r = e / n, c = e - (e / n) * n;
            inv_out[(size_t)i * n * n + (size_t)r * n + s_perm[c]] = A[e];
        }
    for (int lf = 0; lf < 3; ++lf) {
        __syncthreads();
        for (int e = tid; e < n * nq; e += nt) {
            const int r = e / nq, c = e - (e / nq) * nq;
            T[e] = face_coupling_entry(a, i, lf, r / P, c / a.Q, r % P, c % a.Q);
        }
        __syncthreads();
        for (int e = tid; e < n * nq; e += nt) {
            const int r = e / nq, c = e - (e / nq) * nq;
            double s = 0.0;
            for (int k = 0; k < n; ++k) s = fma(A[r * n + k], T[s_perm[k] * nq + c], s);
            aa_trace[(((size_t)i * 3 + lf) * n + r) * nq + c] = s;
        }
    }
}

// BA_SOL [N][3][3Q][7P]
__global__ void k_ba_sol(HdgArgs a, double* __restrict__ ba)
{
    const int n = 7 * a.P, nq = 3 * a.Q;
    const int64_t total = (int64_t)a.N * 3 * nq * n;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
         g += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(g % n);
        const int r = (int)((g / n) % nq);
        const int lf = (int)((g / ((int64_t)n * nq)) % 3);
        const int i = (int)(g / ((int64_t)n * nq * 3));
        ba[g] = ba_sol_entry(a, i, lf, r / a.Q, c / a.P, r % a.Q, c % a.P);
    }
}

// One thread per entry of every block of K (block rows by face, the
// pattern of bsr_ptr / bsr_col): the sum of the diagonal int_fc_fc block
// and -BA_SOL[t, l] AA_TRACE[t, il] of each side whose element lists the
// column face as its local face il.
__global__ void k_assemble_trace(HdgArgs a, const int32_t* __restrict__ t1, const int32_t* __restrict__ l1,
                                 const int32_t* __restrict__ t2, const int32_t* __restrict__ l2,
                                 const uint8_t* __restrict__ mirror, const int32_t* __restrict__ tri_faces,
                                 const int32_t* __restrict__ bsr_row, const int32_t* __restrict__ bsr_col,
                                 int64_t nnzb, const double* __restrict__ aa_trace,
                                 const double* __restrict__ ba, double* __restrict__ val,
                                 unsigned long long* __restrict__ nnz)
{
    const int n = 7 * a.P, nq = 3 * a.Q, Q = a.Q;
    const int64_t bb = (int64_t)nq * nq;
    const int64_t total = nnzb * bb;
    unsigned long long cnt = 0;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
         g += (int64_t)gridDim.x * blockDim.x) {
        const int64_t blk = g / bb;
        const int e = (int)(g - blk * bb);
        const int r = e / nq, c = e - (e / nq) * nq;
        const int f = bsr_row[blk], col = bsr_col[blk];
        double v = 0.0;
        if (col == f && r / Q == c / Q)
            v = a.fcfc[((size_t)(r % Q) * Q + (c % Q)) * a.F + f];
        for (int side = 0; side < 2; ++side) {
            const int t = side == 0 ? t1[f] : t2[f];
            if (t < 0) continue;
            const int l = side == 0 ? l1[f] : l2[f];
            // the partner side of a reversed periodic pair reads its rows mirrored
            const int rr = (side == 1 && mirror[f]) ? (r / Q) * Q + (Q - 1 - r % Q) : r;
            const double* brow = ba + (((size_t)t * 3 + l) * nq + rr) * n;
            for (int il = 0; il < 3; ++il) {
                if (tri_faces[t * 3 + il] != col) continue;
                const double* at = aa_trace + ((size_t)t * 3 + il) * n * nq + c;
                double d = 0.0;
                for (int k = 0; k < n; ++k) d = fma(brow[k], at[(size_t)k * nq], d);
                v -= d;
            }
        }
        if (fabs(v) <= a.drop_tol) v = 0.0;
        val[g] = v;
        cnt += v != 0.0;
    }
    // integer sum: the count does not depend on the order
    for (int m = 32; m > 0; m >>= 1) cnt += __shfl_xor(cnt, m, 64);
    if ((threadIdx.x & 63) == 0 && cnt) atomicAdd(nnz, cnt);
}

__global__ void k_row_of_block(const int32_t* __restrict__ ptr, int F, int32_t* __restrict__ row)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    for (int e = ptr[f]; e < ptr[f + 1]; ++e) row[e] = f;
}

inline unsigned grid_for(int64_t n, int bs) { return (unsigned)std::min<int64_t>((n + bs - 1) / bs, 1 << 20); }

template <typename Fn>
int hguard(char* err, int errlen, Fn&& fn)
{
    try {
        return fn();
    } catch (const std::string& e) {
        if (err && errlen > 0) {
            size_t k = 0;
            for (; k + 1 < (size_t)errlen && k < e.size(); ++k) err[k] = e[k];
            err[k] = 0;
        }
        return -2;
    }
}

#define HD_TRY(expr)                                                        \
    do {                                                                    \
        hipError_t e_ = (expr);                                             \
        if (e_ != hipSuccess) {                                             \
            throw std::string(#expr " failed: ") + hipGetErrorString(e_);  \
        }                                                                   \
    } while (0)

}  // namespace

extern "C" int pbte_hdg_local(const pbte_hdg_desc* d, double* inv_aa_sol, double* aa_trace, double* ba_sol,
                              void* stream, char* err, int32_t errlen)
{
    return hguard(err, errlen, [&] {
        if (!d || d->nd < 1 || d->nd > 15 || d->nf < 2 || d->n_tris < 1)
            throw std::string("hdg_local: unsupported sizes");
        const HdgArgs a = make_args(d);
        hipStream_t st = static_cast<hipStream_t>(stream);
        const int n = 7 * a.P, nq = 3 * a.Q;
        const size_t lds = (size_t)(n * n + n * nq) * sizeof(double);
        if (lds > 65536)
            HD_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(k_local_ops),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        int32_t* nsing = nullptr;
        HD_TRY(hipMalloc(&nsing, sizeof(int32_t)));
        HD_TRY(hipMemsetAsync(nsing, 0, sizeof(int32_t), st));
        hipLaunchKernelGGL(k_local_ops, dim3(a.N), dim3(256), lds, st, a, inv_aa_sol, aa_trace, nsing);
        hipLaunchKernelGGL(k_ba_sol, dim3(grid_for((int64_t)a.N * 3 * nq * n, 256)), dim3(256), 0, st, a, ba_sol);
        int32_t h = 0;
        hipError_t e1 = hipGetLastError();
        hipError_t e2 = hipMemcpyAsync(&h, nsing, sizeof(int32_t), hipMemcpyDeviceToHost, st);
        hipError_t e3 = hipStreamSynchronize(st);
        (void)hipFree(nsing);
        HD_TRY(e1);
        HD_TRY(e2);
        HD_TRY(e3);
        if (h > 0) throw std::string("hdg_local: ") + std::to_string(h) + " singular local operators";
        return 0;
    });
}

namespace {
__global__ void k_zero_u64(unsigned long long* p)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0ull;
}

// the assembly's launches: counter cleared (a kernel, which a captured
// graph may hold), the block-row index of every block, the block values
void launch_assemble(const HdgArgs& a, const int32_t* t1, const int32_t* l1, const int32_t* t2, const int32_t* l2,
                     const uint8_t* mirror, const int32_t* tri_faces, const int32_t* bsr_ptr, const int32_t* bsr_col,
                     int64_t nnzb, const double* aa_trace, const double* ba_sol, double* bsr_val, int32_t* row,
                     unsigned long long* cnt, hipStream_t st)
{
    const int nq = 3 * a.Q;
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, cnt);
    hipLaunchKernelGGL(k_row_of_block, dim3((a.F + 255) / 256), dim3(256), 0, st, bsr_ptr, a.F, row);
    hipLaunchKernelGGL(k_assemble_trace, dim3(grid_for(nnzb * nq * nq, 256)), dim3(256), 0, st, a, t1, l1, t2, l2,
                       mirror, tri_faces, row, bsr_col, nnzb, aa_trace, ba_sol, bsr_val, cnt);
}

// Device scratch of the allocation-free assembly, one copy per GPU for the
// life of the process: the nonzero counter at offset 0, the block-row
// indices from offset 256.  Calls that overlap on one GPU share it and race.
constexpr size_t kScratchBytes = 16u << 20;
constexpr int64_t kScratchMaxBlocks = (int64_t)(kScratchBytes - 256) / 4;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

// the scratch's device address on the current device, looked up once per
// device (the array's host-side name is not a device address)
void* scratch_address()
{
    static std::mutex mu;
    static std::vector<void*> addr;
    int dev = 0;
    HD_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    if ((int)addr.size() <= dev) addr.resize(dev + 1, nullptr);
    if (!addr[dev]) {
        void* p = nullptr;
        HD_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        addr[dev] = p;
    }
    return addr[dev];
}

void read_count(unsigned long long* cnt, int64_t* nnz_host, hipStream_t st)
{
    unsigned long long h = 0;
    HD_TRY(hipGetLastError());
    HD_TRY(hipMemcpyAsync(&h, cnt, sizeof(h), hipMemcpyDeviceToHost, st));
    HD_TRY(hipStreamSynchronize(st));
    if (nnz_host) *nnz_host = (int64_t)h;
}
}  // namespace

extern "C" int pbte_hdg_assemble(const pbte_hdg_desc* d, const int32_t* t1, const int32_t* l1, const int32_t* t2,
                                 const int32_t* l2, const uint8_t* mirror, const int32_t* tri_faces,
                                 const int32_t* bsr_ptr, const int32_t* bsr_col, int64_t nnzb,
                                 const double* aa_trace, const double* ba_sol, double* bsr_val, int64_t* nnz_host,
                                 void* stream, char* err, int32_t errlen)
{
    return hguard(err, errlen, [&] {
        const HdgArgs a = make_args(d);
        hipStream_t st = static_cast<hipStream_t>(stream);
        int32_t* row = nullptr;
        unsigned long long* cnt = nullptr;
        HD_TRY(hipMalloc(&row, std::max<int64_t>(nnzb, 1) * sizeof(int32_t)));
        if (hipMalloc(&cnt, sizeof(unsigned long long)) != hipSuccess) {
            (void)hipFree(row);
            throw std::string("hdg_assemble: hipMalloc of the counter failed");
        }
        try {
            launch_assemble(a, t1, l1, t2, l2, mirror, tri_faces, bsr_ptr, bsr_col, nnzb, aa_trace, ba_sol, bsr_val,
                            row, cnt, st);
            read_count(cnt, nnz_host, st);
        } catch (...) {
            (void)hipFree(row);
            (void)hipFree(cnt);
            throw;
        }
        (void)hipFree(row);
        (void)hipFree(cnt);
        return 0;
    });
}

extern "C" int pbte_hdg_assemble_static(const pbte_hdg_desc* d, const int32_t* t1, const int32_t* l1,
                                        const int32_t* t2, const int32_t* l2, const uint8_t* mirror,
                                        const int32_t* tri_faces, const int32_t* bsr_ptr, const int32_t* bsr_col,
                                        int64_t nnzb, const double* aa_trace, const double* ba_sol,
                                        double* bsr_val, int64_t* nnz_host, void* stream, char* err,
                                        int32_t errlen)
{
    return hguard(err, errlen, [&] {
        if (nnzb < 0 || nnzb > kScratchMaxBlocks)
            throw std::string("hdg_assemble_static: ") + std::to_string(nnzb) + " blocks exceed the " +
                std::to_string(kScratchMaxBlocks) + " the 16 MiB device scratch holds";
        const HdgArgs a = make_args(d);
        hipStream_t st = static_cast<hipStream_t>(stream);
        unsigned char* base = static_cast<unsigned char*>(scratch_address());
        auto* cnt = reinterpret_cast<unsigned long long*>(base);
        auto* row = reinterpret_cast<int32_t*>(base + 256);
        launch_assemble(a, t1, l1, t2, l2, mirror, tri_faces, bsr_ptr, bsr_col, nnzb, aa_trace, ba_sol, bsr_val, row,
                        cnt, st);
        read_count(cnt, nnz_host, st);
        return 0;
    });
}
