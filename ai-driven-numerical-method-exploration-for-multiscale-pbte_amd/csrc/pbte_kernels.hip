// pbte_kernels.hip -- CDNA4 (gfx950) kernels of the gray linear Callaway DG
// discrete-ordinates solve and the C-ABI that drives them.
//
// Data layout in HBM
//   vdf        [D][N][P]   the reference's layout (sweep.py:132-158); a
//                          cell's P values are contiguous, so the sweep
//                          writes are 12-120 B runs and the moment reduction
//                          reads coalesced rows.
//   element inverse per bundle [k / V][item][V], V = 16 B / sizeof(T),
//                          item = slot * B + direction-in-bundle: the items of
//                          one level are consecutive, so a level's inverses
//                          stream as 64-lane coalesced 16-byte loads.
//   slot record per schedule class [slot] int4: cell | (LDS pool slot + 1)
//                          << 21 and three face words saying where each
//                          upwind trace lives (LDS slot, global value, wall,
//                          nothing); read by every direction of every bundle
//                          of the class.
//   cell record [N][rs]    U, VX, VY (the source, rewritten every sweep), the
//                          face normals times the face length, info words;
//                          one 128-byte line per cell in binary32 at P = 6.
//   CellGeo, face blocks   the binary64 tensors of the residual, moment and
//                          on-the-fly kernels.
//
// A bundle is a set of directions with the same inflow sign on every
// interior and periodic face; such directions have the same reference
// order and the same level schedule, so one workgroup sweeps all of them
// together and loads each cell's direction-independent data once (one
// direction per bundle by default).  The bundles of one direction group
// (the same order and inflow signs) and one size form a schedule class:
// their levels, slot records and wall lists are built and stored once.
#include "ref_element.h"
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <functional>
#include <numeric>
#include <queue>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "pbte_hip.h"

int pbte_detail_levels(int n, const int32_t* order, const int32_t* nb, const int32_t* tri_faces,
                       const int32_t* face_bc, const int32_t* bc_type, const int32_t* periodic_tri,
                       const double* nx, const double* ny, double cx, double cy,
                       std::vector<int32_t>& level_of);

namespace {

// --------------------------------------------------------------------------
// element node numbering: node (i/deg, j/deg) has index j(deg+1) - j(j-1)/2 + i
// local face il runs from vertex il to vertex il+1 of (0,0),(1,0),(0,1)
// --------------------------------------------------------------------------
__host__ __device__ constexpr int node_id(int deg, int i, int j)
{
    return j * (deg + 1) - j * (j - 1) / 2 + i;
}
__host__ __device__ constexpr int face_node(int deg, int il, int a)
{
    return il == 0 ? node_id(deg, a, 0) : (il == 1 ? node_id(deg, deg - a, a) : node_id(deg, 0, deg - a));
}
template <int P> struct Deg;
template <> struct Deg<3> { static constexpr int v = 1; };
template <> struct Deg<6> { static constexpr int v = 2; };
template <> struct Deg<10> { static constexpr int v = 3; };
template <> struct Deg<15> { static constexpr int v = 4; };

// padded per-item / per-cell sizes so that every stream is read with
// 16-byte vector loads
__host__ __device__ constexpr int pad4(int n) { return (n + 3) & ~3; }
template <int P> struct Lay {
    static constexpr int Q = Deg<P>::v + 1;
    static constexpr int SRC = pad4(3 * P);  // source values per cell
    static constexpr int KQ = pad4(Q * Q);   // upwind face block per (cell, face)
    static constexpr int LU = pad4(P * P);   // factors per (slot, direction)
    // cell record of the sweep, one row per cell: the source (U, VX, VY),
    // the outward normals of faces 0-2 scaled by the face length, and the
    // face info words.  Every interior, periodic and wall face block of a
    // straight-sided element is its length times one reversed edge mass
    // matrix, so no per-face block is stored.
    static constexpr int R_N = 3 * P, R_INFO = 3 * P + 6;
    static constexpr int CR = pad4(3 * P + 9);
    static constexpr int CRL = pad4(3 * P + 6);  // the part every item loads
    // the fused binary32 sweep's element scalar area2 / tau_c follows the
    // info words (in the row's padding); that sweep loads the row up to it
    static constexpr int R_FUSE = 3 * P + 9;
    static constexpr int CRF = pad4(3 * P + 10);
};

// row stride of the cell records: a multiple of 128 bytes, so a record
// spans the fewest cache lines (one for P <= 6 in binary32)
template <int P, typename T> constexpr int rec_stride()
{
    return (int)(((size_t)Lay<P>::CR * sizeof(T) + 127) / 128 * 128 / sizeof(T));
}
__host__ inline int rec_stride_rt(int P, size_t sz)
{
    return (int)(((size_t)pad4(3 * P + 9) * sz + 127) / 128 * 128 / sz);
}
static_assert(rec_stride<3, float>() >= Lay<3>::CRF && rec_stride<6, float>() >= Lay<6>::CRF &&
                  rec_stride<10, float>() >= Lay<10>::CRF && rec_stride<15, float>() >= Lay<15>::CRF &&
                  rec_stride<3, double>() >= Lay<3>::CRF && rec_stride<6, double>() >= Lay<6>::CRF &&
                  rec_stride<10, double>() >= Lay<10>::CRF && rec_stride<15, double>() >= Lay<15>::CRF,
              "the fused element scalars fit the padding of every cell record");

// LU entry k of item `item` (slot * B + b) of a bundle with NB items, laid
// out [k / V][item][V] with V = 16 / sizeof(T): lanes holding consecutive
// items read consecutive 16-byte words
template <typename T>
__host__ __device__ constexpr size_t lu_idx(int k, size_t item, size_t NB)
{
    return ((size_t)(k / (16 / (int)sizeof(T))) * NB + item) * (16 / sizeof(T)) + (size_t)(k % (16 / (int)sizeof(T)));
}

// face codes of CellGeo::info
enum { FACE_INTERIOR = 0, FACE_PERIODIC = 1, FACE_ADIABATIC = 2, FACE_THERMAL = 3 };

template <typename T> struct alignas(16) CellGeo {
    T n[6];            // (nx, ny) of local faces 0, 1, 2
    int32_t other[3];  // element across (interior / periodic) or face id (walls)
    int32_t info[3];   // code | (other local face << 2) | (face id << 4)
};

#define HIP_TRY(expr)                                                       \
    do {                                                                    \
        hipError_t e_ = (expr);                                             \
        if (e_ != hipSuccess) {                                             \
            throw std::string(#expr " failed: ") + hipGetErrorString(e_);  \
        }                                                                   \
    } while (0)

template <typename T> T* dalloc(size_t n)
{
    void* p = nullptr;
    if (n == 0) n = 1;
    HIP_TRY(hipMalloc(&p, n * sizeof(T)));
    return static_cast<T*>(p);
}

}  // namespace

struct pbte_ctx {
    pbte_dims dm{};
    int P = 0, Q = 0, N = 0, D = 0, F = 0, deg = 0;
    hipStream_t stream = nullptr;
    std::string err;
    double cv = 1, vg = 1, tau_r = 1, tau_n = 1, tau_c = 1, pi = M_PI;
    int block = 256;            // workgroup size of the plain level kernel
    int items = 256;            // (cell, direction) items per workgroup and level

    // direction tables
    double *cxv = nullptr, *cyv = nullptr, *domega = nullptr;
    float *cxv32 = nullptr, *cyv32 = nullptr;
    // full FP64 tensors (factorisation, residual)
    double *mass = nullptr, *gx = nullptr, *gy = nullptr, *fcm = nullptr, *int_tri = nullptr;
    double* nrm = nullptr;     // [N][3][2] outward normals
    // element data of the factorisation and the transport residual
    CellGeo<double>* geo64 = nullptr;
    double* K64 = nullptr;      // [N][3][KQ] face blocks on the shared face's nodes
    double* bvec64 = nullptr;   // [F][P] tw * ttfc[:, 0] on thermal walls
    int n_bc_th = 0;
    // sweep data: the reversed unit-edge mass matrix Kr, the element across
    // each face, the wall vectors
    double kref64[25] = {};
    float kref32[25] = {};
    int32_t* other = nullptr;   // [N][3]
    int n_wrow = 0;             // wall rows (thermal and adiabatic faces)
    double* wtr64 = nullptr;    // [n_wrow][pad4(P)] wall vectors over the face length
    float* wtr32 = nullptr;
    int n_adw = 0;              // adiabatic wall rows, refreshed every sweep
    int32_t *adw_row = nullptr, *adw_face = nullptr;
    double* adw_rlen = nullptr;
    // adiabatic walls
    int n_wall = 0;
    int32_t *wall_face = nullptr, *wall_tri = nullptr, *wall_lfc = nullptr;
    // periodic faces (Krylov state)
    int n_per = 0;
    int32_t *per_face = nullptr, *per_tri = nullptr;
    // schedule
    int nb = 0;
    int64_t total_levels = 0, max_width = 0;
    int n_classes = 0;          // schedule classes (bundles sharing levels and records)
    int32_t *block_bundle = nullptr, *bundle_ndir = nullptr, *bundle_dir_off = nullptr,
            *bundle_dirs = nullptr, *slot_cell = nullptr, *lev_off = nullptr, *lev_ptr = nullptr;
    int64_t *lu_off = nullptr, *item_off = nullptr;
    int4* irec = nullptr;       // per (class, slot): {cell | pool slot, face words}
    int4* wcell = nullptr;      // per class: inflow wall elements {cell, wall row of faces 0-2}
    int32_t* wc_off = nullptr;
    int max_wall_in = 0;        // most inflow wall elements of one bundle
    double* crec64 = nullptr;   // cell records of the sweep
    float* crec32 = nullptr;
    int rs64 = 0, rs32 = 0;     // their row strides
    uint8_t* bundle_global = nullptr;
    int pool_slots = 0;         // LDS trace slots per workgroup
    size_t lds_blk = 65536;     // LDS bytes one workgroup may take
    int depth = 3;              // levels between a gather and its solve, plus one
    int64_t pool_peak = 0;      // most slots one bundle holds at once
    int64_t global_bundles = 0;
    int variant = 0;            // 0 pipelined level sweep, 1 plain level sweep
    int max_bundle_dirs = 1;
    // upwind reads by kind: previous-sweep values across non-periodic faces
    // (broken cycles), previous-sweep values across periodic faces, every
    // periodic face; and of each the reads whose class is not staged
    int64_t np_old_reads = 0, per_old_reads = 0, periodic_reads = 0;
    int64_t np_old_unstaged = 0, per_old_unstaged = 0, per_unstaged = 0;
    int32_t* stage_list = nullptr;  // per class: staged elements, then staged periodic faces
    int4* stage_off = nullptr;      // per bundle: {start, elements, non-periodic ones, periodic faces}
    int stage_slots[2] = {0, 0};    // most stage slots of a bundle without / with an explicit periodic buffer
    bool staging = true;            // pipelined launches stage those reads where their LDS fits
    std::vector<int64_t> h_lu_off;
    int64_t lu_elems = 0, item_elems = 0;
    // factors
    double* lu64 = nullptr;
    float* lu32 = nullptr;
    bool have64 = false, have32 = false;
    // matrix-free element operators (the refined binary64 sweep)
    double* refop = nullptr;    // [6][P][P]
    double* cellop = nullptr;   // [N][8]
    // the fused binary32 sweep: the reference-element matrices [3][P][P] and
    // the unit-edge mass matrix [Q][Q] of this context (fused_ok once the face
    // tensors are checked, fused_tables once these equal the compiled
    // tables), on when fused32 is set; face lengths [N][3]
    std::vector<double> h_fref;
    std::vector<double> h_flen;
    bool fused_ok = false, fused32 = false, have_cellops = false, fused_tables = false, fused_geometry = false;
    bool split = false;         // fused sweeps at P <= 6 in the split kernel where its LDS fits
    // scratch
    double* src64 = nullptr;   // source of the transport residual
    double* ssraw = nullptr;
    double* red = nullptr;     // reduction partials
    size_t red_len = 0;
    double* small = nullptr;   // small device vector (GMRES coefficients)
    int32_t* counter = nullptr;
    double* spill_row = nullptr;   // 16 values written by idle sweep threads
    // sweep-kernel timing events (armed by pbte_timing)
    std::vector<hipEvent_t> events;
    int ev_used = 0;
    double* hbuf = nullptr;    // pinned host staging
    size_t hbuf_len = 0;
    // Gram-Schmidt coefficients read back behind an event (pbte_vec_cgs2_begin/end)
    double* cg_host = nullptr;
    hipEvent_t cg_event = nullptr;
    int cg_m = 0;

    // GSIS data
    bool gsis = false;
    int gs_nw = 0, gs_nth = 0, gs_variant_b = 0, gs_use_hmin = 1;
    double gs_tc5 = 0, gs_tau_thr = 1, gs_drift = 0;
    // element operators, K = 3P + 9Q outputs per element: gs_op [N][4P][K]
    // (stress work -> AA_SRC rows of T, qx, qy and the per-side trace sums),
    // gs_aat [N][3][3Q][3P] (AA_TRACE rows of T, qx, qy), their binary32
    // copies (made by the first binary32 step), the rows 3P .. 7P for the
    // full UQ (gs_op_hi [N][4P][4P], gs_aat_hi [N][3][3Q][4P]) and the
    // moment preconditioner's gs_op_pc [N][3P][K]
    double *gs_op = nullptr, *gs_aat = nullptr, *gs_op_hi = nullptr, *gs_aat_hi = nullptr, *gs_op_pc = nullptr;
    float *gs_op32 = nullptr, *gs_aat32 = nullptr;
    double *gs_hw = nullptr, *gs_wtfc = nullptr, *gs_wnx = nullptr, *gs_wny = nullptr, *gs_hmin = nullptr,
           *gs_wmom = nullptr, *gs_side = nullptr;
    int32_t *gs_t1 = nullptr, *gs_l1 = nullptr, *gs_t2 = nullptr, *gs_l2 = nullptr, *gs_wface = nullptr,
            *gs_wtri = nullptr, *gs_tri_faces = nullptr;
    uint8_t* gs_flip = nullptr;

    std::vector<void*> owned;
    template <typename T> T* alloc(size_t n)
    {
        T* p = dalloc<T>(n);
        owned.push_back(p);
        return p;
    }
    template <typename T> T* upload(const T* h, size_t n)
    {
        T* p = alloc<T>(n);
        HIP_TRY(hipMemcpy(p, h, n * sizeof(T), hipMemcpyHostToDevice));
        return p;
    }
    ~pbte_ctx()
    {
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
        if (cg_event) (void)hipEventDestroy(cg_event);
        for (void* p : owned) (void)hipFree(p);
        if (hbuf) (void)hipHostFree(hbuf);
        if (cg_host) (void)hipHostFree(cg_host);
    }
    double* host_staging(size_t n)
    {
        if (n > hbuf_len) {
            if (hbuf) HIP_TRY(hipHostFree(hbuf));
            hbuf = nullptr;
            HIP_TRY(hipHostMalloc((void**)&hbuf, n * sizeof(double), hipHostMallocDefault));
            hbuf_len = n;
        }
        return hbuf;
    }
};

namespace {

// ==========================================================================
// element operators: build A_SOL (_kernels.py:106-125), factorise it with
// DGETF2's pivoting (_kernels.py:43-76) and form its inverse from the
// factors, all in binary64; the sweep layout stores the inverse.
// ==========================================================================
struct FactorArgs {
    const int32_t *slot_cell, *bundle_ndir, *bundle_dir_off, *bundle_dirs;
    const int64_t *lu_off, *item_off;
    const double *cxv, *cyv, *mass, *gx, *gy, *fcm;
    const double *refop, *cellop;      // reference-element matrices, 8 scalars per cell
    const CellGeo<double>* geo;
    double inv_tau_c;
    int N;
    double* lu64;
    float* lu32;
    int32_t* singular;
};

// A_SOL of element i for the direction (cx, cy) (_kernels.py:106-125) and
// its DGETF2 factorisation with the reference's pivoting (_kernels.py:43-76),
// in binary64; the diagonal holds reciprocals (the solve multiplies by
// them).  Returns 0 or the first zero pivot (1-based), like INFO.
// The entries (m, l) of the element's tensors come from get(m, l, E):
// E[0] mass, E[1] / E[2] the x / y streaming terms, E[3..5] the three
// face terms.
struct ElemEntries {
    double e[6];
};

template <int P, typename Get>
__device__ __forceinline__ int element_lu_from(Get get, const CellGeo<double>& g, double inv_tau_c, double cx,
                                               double cy, double (&A)[P][P], uint64_t& pv)
{
    double w[3];
#pragma unroll
    for (int il = 0; il < 3; ++il) {
        double sp = cx * g.n[2 * il] + cy * g.n[2 * il + 1];
        w[il] = 0.5 * (sp + fabs(sp));
    }
#pragma unroll
    for (int l = 0; l < P; ++l)
#pragma unroll
        for (int m = 0; m < P; ++m) {
            const ElemEntries E = get(m, l);
            double v = (inv_tau_c * E.e[0] - cx * E.e[1] - cy * E.e[2]);
            v = v + w[0] * E.e[3];
            v = v + w[1] * E.e[4];
            v = v + w[2] * E.e[5];
            A[m][l] = v;
        }
    pv = 0;
    int info = 0;
#pragma unroll
    for (int k = 0; k < P; ++k) {
        int p = k;
        double big = fabs(A[k][k]);
#pragma unroll
        for (int r = k + 1; r < P; ++r) {
            double v = fabs(A[r][k]);
            if (v > big) { big = v; p = r; }
        }
        pv |= (uint64_t)p << (4 * k);
#pragma unroll
        for (int r = k + 1; r < P; ++r) {
            if (p == r) {
#pragma unroll
                for (int j = 0; j < P; ++j) {
                    double t = A[k][j]; A[k][j] = A[r][j]; A[r][j] = t;
                }
            }
        }
        if (A[k][k] == 0.0) {
            if (info == 0) info = k + 1;
            continue;
        }
        const double rcp = 1.0 / A[k][k];
#pragma unroll
        for (int r = k + 1; r < P; ++r) A[r][k] = A[r][k] * rcp;
#pragma unroll
        for (int r = k + 1; r < P; ++r) {
            const double ark = A[r][k];
#pragma unroll
            for (int j = k + 1; j < P; ++j) A[r][j] -= ark * A[k][j];
        }
        A[k][k] = rcp;
    }
    return info;
}

// element_lu_from on the element tensors as stored: mass [P][P][N],
// streaming [N][P][P], faces [N][3][P][P]
template <int P>
__device__ __forceinline__ int element_lu(const double* __restrict__ mass, const double* __restrict__ gx,
                                          const double* __restrict__ gy, const double* __restrict__ fcm,
                                          const CellGeo<double>& g, double inv_tau_c, int N, int i, double cx,
                                          double cy, double (&A)[P][P], uint64_t& pv)
{
    const double* fc = fcm + (size_t)i * 3 * P * P;
    return element_lu_from<P>(
        [&](int m, int l) {
            ElemEntries E;
            E.e[0] = mass[((size_t)m * P + l) * N + i];
            E.e[1] = gx[((size_t)i * P + m) * P + l];
            E.e[2] = gy[((size_t)i * P + m) * P + l];
#pragma unroll
            for (int il = 0; il < 3; ++il) E.e[3 + il] = fc[(il * P + m) * P + l];
            return E;
        },
        g, inv_tau_c, cx, cy, A, pv);
}

// element_lu_from on tensors rebuilt from the reference-element matrices
// ref [6][P][P] (mass, d/dx, d/dy, faces 1-3) and the 8 scalars of the cell
// co (2|T|, the four inverse-Jacobian terms, the three scaled face
// lengths), in integration.py's products one rounding at a time (separate
// statements: no fma), so every entry equals the stored tensor's bit for bit
template <int P>
__device__ __forceinline__ int element_lu_ref(const double* ref, const double (&co)[8], const CellGeo<double>& g,
                                              double inv_tau_c, double cx, double cy, double (&A)[P][P],
                                              uint64_t& pv)
{
    constexpr int PP = P * P;
    return element_lu_from<P>(
        [&](int m, int l) {
            const int e = m * P + l;
            ElemEntries E;
            E.e[0] = ref[e] * co[0];
            const double x1 = co[1] * ref[PP + e];
            const double x2 = co[2] * ref[2 * PP + e];
            const double xs = x1 + x2;
            E.e[1] = xs * co[0];
            const double y1 = co[3] * ref[PP + e];
            const double y2 = co[4] * ref[2 * PP + e];
            const double ys = y1 + y2;
            E.e[2] = ys * co[0];
#pragma unroll
            for (int il = 0; il < 3; ++il) E.e[3 + il] = co[5 + il] * ref[(3 + il) * PP + e];
            return E;
        },
        g, inv_tau_c, cx, cy, A, pv);
}

// v = inv(A) v from element_lu's factors, DGETRS's loop order
// (_kernels.py:78-101): row interchanges, unit lower, then upper
template <int P>
__device__ __forceinline__ void element_lu_solve(const double (&A)[P][P], uint64_t pv, double (&v)[P])
{
#pragma unroll
    for (int k = 0; k < P; ++k) {
        const int p = (int)((pv >> (4 * k)) & 15u);
#pragma unroll
        for (int r = k + 1; r < P; ++r)
            if (p == r) { double t = v[k]; v[k] = v[r]; v[r] = t; }
    }
#pragma unroll
    for (int k = 0; k < P; ++k)
#pragma unroll
        for (int r = k + 1; r < P; ++r) v[r] -= v[k] * A[r][k];
#pragma unroll
    for (int k = P - 1; k >= 0; --k) {
        v[k] = v[k] * A[k][k];
#pragma unroll
        for (int r = 0; r < k; ++r) v[r] -= v[k] * A[r][k];
    }
}

template <int P>
__global__ void __launch_bounds__(256) factor_kernel(FactorArgs a)
{
    const int bundle = blockIdx.y;
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.N) return;
    const int B = a.bundle_ndir[bundle];
    const int i = a.slot_cell[a.item_off[bundle] + s];
    const int N = a.N;
    const CellGeo<double> g = a.geo[i];
    for (int b = 0; b < B; ++b) {
        const int d = a.bundle_dirs[a.bundle_dir_off[bundle] + b];
        double A[P][P];
        uint64_t pv;
        if (element_lu<P>(a.mass, a.gx, a.gy, a.fcm, g, a.inv_tau_c, N, i, a.cxv[d], a.cyv[d], A, pv))
            atomicAdd(a.singular, 1);
        // the inverse, column by column from the pivoted factors (DGETRS on
        // the unit vectors), each column stored as it is solved (no P x P
        // array beside the factors: it spilled at P = 10): the sweep
        // applies x = inv(A) b
        const int64_t base = a.lu_off[bundle];
        const size_t item = (size_t)s * B + b, NB = (size_t)N * B;
#pragma unroll 1
        for (int j = 0; j < P; ++j) {
            double v[P];
#pragma unroll
            for (int k = 0; k < P; ++k) v[k] = k == j ? 1.0 : 0.0;
            element_lu_solve<P>(A, pv, v);
#pragma unroll
            for (int r = 0; r < P; ++r) {
                if (a.lu64) a.lu64[base + lu_idx<double>(r * P + j, item, NB)] = v[r];
                if (a.lu32) a.lu32[base + lu_idx<float>(r * P + j, item, NB)] = (float)v[r];
            }
        }
    }
}

// The same factorisation and inverse, one wave per 64 consecutive items of
// a bundle (item = slot * B + direction), the operators rebuilt from the
// reference-element matrices (LDS) and 8 scalars per cell (bit for bit the
// stored tensors), each item's inverse staged in LDS as it is solved and
// then written as whole 16-byte words, word plane by word plane, so a wave
// stores 1 KiB runs; binary64 and binary32 copies from one pass when both
// are asked for.  Grid: (ceil(N * max B / 64), bundles).
template <int P>
__global__ void __launch_bounds__(64) factor_ref_kernel(FactorArgs a)
{
    constexpr int PP = P * P;
    constexpr int PV = pad4(PP);
    __shared__ double ref[6 * PP];
    __shared__ double inv[PP][64];
    const int bundle = blockIdx.y;
    const int t = threadIdx.x;
    for (int k = t; k < 6 * PP; k += 64) ref[k] = a.refop[k];
    __syncthreads();
    const int B = a.bundle_ndir[bundle];
    const size_t NB = (size_t)a.N * B;
    const size_t item = (size_t)blockIdx.x * 64 + t;
    if ((size_t)blockIdx.x * 64 >= NB) return;          // uniform over the block
    const bool valid = item < NB;
    if (valid) {
        const int s = (int)(item / B), b = (int)(item % B);
        const int i = a.slot_cell[a.item_off[bundle] + s];
        const int d = a.bundle_dirs[a.bundle_dir_off[bundle] + b];
        double co[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) co[k] = a.cellop[(size_t)i * 8 + k];
        const CellGeo<double> g = a.geo[i];
        double A[P][P];
        uint64_t pv;
        if (element_lu_ref<P>(ref, co, g, a.inv_tau_c, a.cxv[d], a.cyv[d], A, pv)) atomicAdd(a.singular, 1);
#pragma unroll 1
        for (int j = 0; j < P; ++j) {
            double v[P];
#pragma unroll
            for (int k = 0; k < P; ++k) v[k] = k == j ? 1.0 : 0.0;
            element_lu_solve<P>(A, pv, v);
#pragma unroll
            for (int r = 0; r < P; ++r) inv[r * P + j][t] = v[r];
        }
    }
    __syncthreads();
    if (!valid) return;
    const int64_t base = a.lu_off[bundle];
    if (a.lu64) {
        double2* o = reinterpret_cast<double2*>(a.lu64 + base);
#pragma unroll
        for (int w = 0; w < PV / 2; ++w) {
            double2 x;
            x.x = 2 * w < PP ? inv[2 * w][t] : 0.0;
            x.y = 2 * w + 1 < PP ? inv[2 * w + 1][t] : 0.0;
            o[(size_t)w * NB + item] = x;
        }
    }
    if (a.lu32) {
        float4* o = reinterpret_cast<float4*>(a.lu32 + base);
#pragma unroll
        for (int w = 0; w < PV / 4; ++w) {
            float4 x;
            x.x = 4 * w < PP ? (float)inv[4 * w][t] : 0.0f;
            x.y = 4 * w + 1 < PP ? (float)inv[4 * w + 1][t] : 0.0f;
            x.z = 4 * w + 2 < PP ? (float)inv[4 * w + 2][t] : 0.0f;
            x.w = 4 * w + 3 < PP ? (float)inv[4 * w + 3][t] : 0.0f;
            o[(size_t)w * NB + item] = x;
        }
    }
}

// ==========================================================================
// source vectors: src = U + cx VX + cy VY per cell (the Callaway scattering
// source of _kernels.py:153-167 split by its dependence on the direction)
// ==========================================================================
template <int P, typename T>
__global__ void __launch_bounds__(256) source_kernel(const double* __restrict__ mass,
                                                     const double* __restrict__ ts,
                                                     const double* __restrict__ qxs,
                                                     const double* __restrict__ qys, T* __restrict__ src,
                                                     int stride, int N, double cv, double pi, double tau_r,
                                                     double tau_n, double vg)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= N * P) return;
    const int i = g / P, m = g - (g / P) * P;
    double u = 0.0, vx = 0.0, vy = 0.0;
#pragma unroll
    for (int l = 0; l < P; ++l) {
        const double mlm = mass[((size_t)l * P + m) * N + i];
        const double t = ts[(size_t)l * N + i];
        const double al = cv * t / 4.0 / pi / tau_r + 1.0 / tau_n * (cv * t / 4.0 / pi);
        const double bx = 1.0 / tau_n * (3.0 / 4.0 / pi * qxs[(size_t)l * N + i] / vg / vg);
        const double by = 1.0 / tau_n * (3.0 / 4.0 / pi * qys[(size_t)l * N + i] / vg / vg);
        u += mlm * al;
        vx += mlm * bx;
        vy += mlm * by;
    }
    T* o = src + (size_t)i * stride;
    o[m] = (T)u;
    o[P + m] = (T)vx;
    o[2 * P + m] = (T)vy;
}

// ==========================================================================
// the level-set sweep (_kernels.py:217-239 with the inflow of 169-215 and
// the triangular solve of 78-101).  One workgroup per direction bundle;
// thread t owns direction t % B of the bundle for the whole sweep and slot
// t / B of every level; a barrier separates levels.
//
// Per (cell, direction) item:
//   x = U + cx VX + cy VY + sum over inflow wall faces of w v    the source
//       and wall inflow (_kernels.py:153-167, 202-215)
//   for each face il, in face order, with w = 0.5 (s - |s|), s = c . n L:
//       x[node(il, r)] -= w (Kr u)[r]            the upwind inflow (_kernels.py:183-201)
//   x = inv(A) x                                 (_kernels.py:78-101)
// Kr is the unit-edge mass matrix with its columns reversed (the element
// across numbers the shared edge's nodes the other way round), L the face
// length, u the upwind element's values at its face nodes (zero on faces
// without upwind inflow) and v a wall's P-vector over L: -tw ttfc[:, 0] for
// a thermal wall, flux_wall for an adiabatic one.  L Kr is the reference's
// face block int_tri_tri_fc to rounding (checked when the context is
// created).
// ==========================================================================
template <typename T> struct SweepArgs {
    const int32_t *block_bundle, *bundle_ndir, *bundle_dir_off, *bundle_dirs, *lev_off, *lev_ptr;
    const int64_t *lu_off, *item_off;
    const T *lu;
    const T *cxv, *cyv;
    const T* crec;              // cell records, row stride rs
    const int32_t* other;       // [N][3] element across each face (plain kernel)
    const T* wtr;               // wall vectors over L, [wall row][PV]
    const int4* wcell;          // per class: its inflow wall elements {cell, row of faces 0-2}
    const int32_t* wc_off;
    const T* pbuf;
    T* vdf;
    const int4* irec;           // per (class, slot): {cell | pool slot, face words}
    T* spill_row;               // write target of idle threads (P values)
    const uint8_t* bundle_global;
    const int32_t* stage_list;  // per class: its staged reads, (element or face) << 2 | face across
    const int4* stage_off;      // per bundle: {its class's start in stage_list, previous-sweep reads
                                // across non-periodic faces, across periodic faces, periodic reads}
    int N, F, rs, use_pbuf, homogeneous;
    int stage_on;               // the stage slots follow the spill slot in LDS
    int lane_zero;              // 0 at run time; makes a uniform load per-lane
    int pool;                   // LDS pool slots
    int wbase;                  // LDS offset (in values) of the wall inflow vectors
    int fbase;                  // LDS offset (in values) of the split sweep's factor buffers
    int nthr;                   // threads that sweep (the context's workgroup size; a batch
                                // launches the widest of its contexts)
    T kref[25];                 // Kr, Q x Q, row major
    // the on-the-fly sweep: element tensors and binary64 directions
    const double *mass, *gx, *gy, *fcm, *cxv64, *cyv64;
    const CellGeo<double>* geo;
    double inv_tau_c;
    // the refined binary64 sweep: reference-element matrices [6][P][P]
    // (mass, d/dxi, d/deta, the three edge matrices), per-cell scalars
    // [N][8] (twice the area, the four gradient coefficients, the three
    // edge lengths as the edge tensors scale them), the stored binary32
    // inverse of the same schedule
    const double *refop, *cellop;
    const float* lu32;
};

// Face word of an item record: code in bits 28-31, the local face of the
// element across in bits 26-27, a value in bits 0-25.  What an inflow face
// reads:
//   UP_LDS / PER_LDS      this sweep's trace of the upwind element (periodic
//                         partner) from its LDS pool slot (value)
//   ADIABATIC / THERMAL   a wall: the element's summed wall inflow (value:
//                         the element's index in the bundle's wall list)
//   OLD / PER_OLD         previous-sweep values of an element the reference
//                         order visits later (broken cycle), global memory
//                         (value: the element)
//   UP_GLOBAL/PER_GLOBAL  this sweep's values from global memory (value: the
//                         element), only in a bundle whose LDS pool would
//                         overflow, which then runs with full barriers
//   OLD_ST / PER_OLD_ST   OLD / PER_OLD with the value staged: the sweep's
//                         prologue copies the trace every such face reads
//                         into an LDS stage slot (value: the stage slot in
//                         bits 0-12; PER_OLD_ST also the face's stage slot
//                         for an explicit periodic buffer in bits 13-25)
//   PER_LDS_ST            PER_LDS (pool slot in bits 0-12) with the face's
//                         stage slot for an explicit periodic buffer in bits
//                         13-25
// With an explicit periodic buffer every periodic face reads that instead
// (a staged one from its periodic stage slot).  A launch whose LDS has no
// room for the stage slots reads the staged faces from global memory (the
// element across from the element's neighbour list).
enum {
    REC_UP_LDS = 0, REC_OLD_ST = 1, REC_ADIABATIC = 2, REC_THERMAL = 3, REC_OLD = 4, REC_PER_LDS = 5,
    REC_PER_OLD_ST = 6, REC_PER_OLD = 7, REC_UP_GLOBAL = 8, REC_PER_GLOBAL = 9, REC_PER_LDS_ST = 10, REC_NONE = 15
};
constexpr int kFaceValBits = 26;
constexpr int kFaceValMask = (1 << kFaceValBits) - 1;
constexpr int kStageBits = 13;
constexpr int kStageMax = 1 << kStageBits;      // stage slots of one class (each list)
constexpr int kStageMask = kStageMax - 1;
__host__ __device__ constexpr int face_word(int code, int ilp, int val)
{
    return (int)(((unsigned)code << 28) | ((unsigned)ilp << kFaceValBits) | (unsigned)val);
}
__host__ __device__ constexpr int fw_code(int w) { return (int)((unsigned)w >> 28); }
__host__ __device__ constexpr int fw_ilp(int w) { return (w >> kFaceValBits) & 3; }
__host__ __device__ constexpr int fw_val(int w) { return w & kFaceValMask; }

// item record word x: the element in bits 0-20, the LDS pool slot it writes
// plus one in bits 21-31 (0: no downwind reader reads it through LDS)
constexpr int kCellBits = 21;
constexpr int kCellMask = (1 << kCellBits) - 1;
constexpr int kPoolMax = (1 << (32 - kCellBits)) - 2;
__host__ __device__ constexpr int rec_cell(int x) { return x & kCellMask; }
__host__ __device__ constexpr int rec_pool(int x) { return (int)((unsigned)x >> kCellBits) - 1; }

// record-code classes as bit masks over the code: one shift and one test,
// no short-circuit control flow
__host__ __device__ constexpr int code_bit(int c) { return 1 << c; }
constexpr int kRecWall = code_bit(REC_ADIABATIC) | code_bit(REC_THERMAL);
constexpr int kRecStaged = code_bit(REC_OLD_ST) | code_bit(REC_PER_OLD_ST) | code_bit(REC_PER_LDS_ST);
constexpr int kRecNeighbour = code_bit(REC_UP_LDS) | code_bit(REC_OLD) | code_bit(REC_PER_LDS) |
                              code_bit(REC_PER_OLD) | code_bit(REC_UP_GLOBAL) | code_bit(REC_PER_GLOBAL) | kRecStaged;
constexpr int kRecPeriodic = code_bit(REC_PER_LDS) | code_bit(REC_PER_OLD) | code_bit(REC_PER_GLOBAL) |
                             code_bit(REC_PER_OLD_ST) | code_bit(REC_PER_LDS_ST);
__device__ __forceinline__ bool code_in(int mask, int code) { return ((mask >> code) & 1) != 0; }

// value m of the direction's source plus the wall inflow v
template <int P, typename T>
__device__ __forceinline__ T source_plus(const T* cr, T cx, T cy, int m, T v)
{
    return cr[m] + cx * cr[P + m] + cy * cr[2 * P + m] + v;
}

// LDS trace layout: per (slot, direction of the bundle) the element's
// values on its three faces, face-major, each face padded to TQ values
template <int P> struct Tr {
    static constexpr int Q = Deg<P>::v + 1;
    static constexpr int TQ = Q == 2 ? 2 : pad4(Q);     // padded trace of one face
    // values per (slot, direction), a multiple of 4 so that every slot and
    // the wall vectors (read from the zero slot) stay 16-byte aligned
    static constexpr int W = pad4(3 * TQ);
};
__host__ inline int trace_pad(int Q) { return Q == 2 ? 2 : pad4(Q); }

// wall entries of the bundle's wall list: the wall row of each face, -1
// when the face is no inflow wall; bit 30 marks a thermal wall (it emits
// nothing in a homogeneous sweep)
constexpr int kWallThermal = 1 << 30;

// the summed wall inflow of element i for one direction, value m:
// sum over its inflow wall faces, in face order, of w v[m]
template <int P, typename T>
__device__ __forceinline__ T wall_inflow(const T* cr, int4 e, T cx, T cy, const T* wtr, T thermal_on, int m)
{
    using L = Lay<P>;
    constexpr int PV = pad4(P);
    const int row[3] = {e.y, e.z, e.w};
    T v = T(0);
#pragma unroll
    for (int il = 0; il < 3; ++il) {
        if (row[il] < 0) continue;
        const T sp = cx * cr[L::R_N + 2 * il] + cy * cr[L::R_N + 2 * il + 1];
        const T w = T(0.5) * (sp - fabs(sp));
        const T on = (row[il] & kWallThermal) ? thermal_on : T(1);
        v += w * (wtr[(size_t)(row[il] & (kWallThermal - 1)) * PV + m] * on);
    }
    return v;
}

// the face coupling of one inflow face: x[node(il, r)] -= w (Kr u)[r], the
// product Kr u accumulated first
template <int P, typename T>
__device__ __forceinline__ void face_update(T (&x)[P], int il, T w, const T* K, const T* u)
{
    constexpr int DEG = Deg<P>::v;
    constexpr int Q = DEG + 1;
#pragma unroll
    for (int r = 0; r < Q; ++r) {
        T t = K[r * Q] * u[0];
#pragma unroll
        for (int q = 1; q < Q; ++q) t += K[r * Q + q] * u[q];
        x[face_node(DEG, il, r)] -= w * t;
    }
}

// x = inv(A) b, each row accumulated in column order
template <int P, typename T, class E>
__device__ __forceinline__ void apply_inverse(T (&x)[P], E&& e)
{
    T y[P];
#pragma unroll
    for (int r = 0; r < P; ++r) {
        T v = e(r * P) * x[0];
#pragma unroll
        for (int c = 1; c < P; ++c) v += e(r * P + c) * x[c];
        y[r] = v;
    }
#pragma unroll
    for (int r = 0; r < P; ++r) x[r] = y[r];
}

__device__ __forceinline__ float recip(float v) { return __builtin_amdgcn_rcpf(v); }
__device__ __forceinline__ double recip(double v) { return 1.0 / v; }

// the fused element solve: A x = b with A built where it is used from the
// reference-element matrices (mass M, D_xi, D_eta and the unit-edge mass
// matrix Kf of a face, compile-time tables whose zero entries drop out) and
// the element's five scalars in its cell record,
//   A = (area2 / tau_c) M - area2 (cx ax + cy ay) D_xi - area2 (cx bx + cy by) D_eta
//       + sum over faces of w+ L Kf on the face's nodes,  w+ = max(c.n, 0)
// (_kernels.py:106-125 builds the same operator from the element tensors,
// which are these products), then factorised without pivoting: A + A^T is
// positive definite (the mass term, plus the upwind boundary term: the
// volume advection term is skew up to the boundary integral), so no leading
// block is singular.  Doolittle elimination with the right-hand side
// carried along, then back substitution.
template <typename T> __device__ __forceinline__ void fma_const(T& acc, double k, T v)
{
    if (k != 0.0) acc += (T)k * v;
}

template <int P> constexpr double unit_edge(int r, int q)
{
    return RefElem<P>::E0[face_node(Deg<P>::v, 0, r) * P + face_node(Deg<P>::v, 0, q)];
}

// A of one item from its element scalar am = area2/tau_c and the
// length-scaled outward normals nl of faces 0, 1, 2 (edges 12, 23, 31).
// The advection coefficients -area2 (c . grad xi) and -area2 (c . grad eta)
// of D_xi and D_eta are the flux speeds c . (n L) through faces 2 and 0:
// area2 grad xi = -(n L) of edge 31, area2 grad eta = -(n L) of edge 12
// (checked per element when the scalars are set)
template <int P, typename T>
__device__ __forceinline__ void fused_build(T (&A)[P][P], T am, const T* nl, T cx, T cy)
{
    using R = RefElem<P>;
    constexpr int DEG = Deg<P>::v;
    constexpr int Q = DEG + 1;
    const T bx = cx * nl[4] + cy * nl[5];
    const T by = cx * nl[0] + cy * nl[1];
#pragma unroll
    for (int m = 0; m < P; ++m)
#pragma unroll
        for (int l = 0; l < P; ++l) {
            T v = T(-0.0);      // -0 + x folds to x
            fma_const(v, R::M[m * P + l], am);
            fma_const(v, R::Dx[m * P + l], bx);
            fma_const(v, R::Dy[m * P + l], by);
            A[m][l] = v;
        }
#pragma unroll
    for (int il = 0; il < 3; ++il) {
        const T sp = cx * nl[2 * il] + cy * nl[2 * il + 1];
        const T w = T(0.5) * (sp + fabs(sp));
#pragma unroll
        for (int r = 0; r < Q; ++r)
#pragma unroll
            for (int q = 0; q < Q; ++q) fma_const(A[face_node(DEG, il, r)][face_node(DEG, il, q)], unit_edge<P>(r, q), w);
    }
}

// Doolittle elimination in place: the multipliers below the diagonal, U
// above it, the reciprocal of U's diagonal on it
template <int P, typename T>
__device__ __forceinline__ void fused_eliminate(T (&A)[P][P])
{
#pragma unroll
    for (int k = 0; k < P; ++k) {
        const T rk = recip(A[k][k]);
#pragma unroll
        for (int i = k + 1; i < P; ++i) {
            const T f = A[i][k] * rk;
#pragma unroll
            for (int j = k + 1; j < P; ++j) A[i][j] -= f * A[k][j];
            A[i][k] = f;
        }
        A[k][k] = rk;
    }
}

// x = A^-1 x from fused_eliminate's factors, read through e(i, j)
template <int P, typename T, class E>
__device__ __forceinline__ void fused_substitute(T (&x)[P], E&& e)
{
#pragma unroll
    for (int k = 0; k < P; ++k)
#pragma unroll
        for (int i = k + 1; i < P; ++i) x[i] -= e(i, k) * x[k];
#pragma unroll
    for (int i = P - 1; i >= 0; --i) {
        T v = x[i];
#pragma unroll
        for (int j = i + 1; j < P; ++j) v -= e(i, j) * x[j];
        x[i] = v * e(i, i);
    }
}

template <int P, typename T>
__device__ __forceinline__ void fused_solve(T (&x)[P], const T* cr, T cx, T cy)
{
    using L = Lay<P>;
    T A[P][P];
    fused_build<P, T>(A, cr[L::R_FUSE], cr + L::R_N, cx, cy);
    fused_eliminate<P, T>(A);
    fused_substitute<P, T>(x, [&](int i, int j) { return A[i][j]; });
}

__device__ __forceinline__ int bits_of(float v) { return __float_as_int(v); }
__device__ __forceinline__ int bits_of(double v) { return (int)(uint32_t)(uint64_t)__double_as_longlong(v); }

// the plain level kernel: every value from global memory, one level at a
// time.  Same per-item arithmetic as the pipelined kernel below (the two
// agree bit for bit); kept as the reference schedule for tests.
template <int P, typename T>
__global__ void __launch_bounds__(256) sweep_kernel(SweepArgs<T> a)
{
    using L = Lay<P>;
    constexpr int DEG = Deg<P>::v;
    constexpr int Q = DEG + 1;
    constexpr int TQ = Tr<P>::TQ;
    const int bundle = a.block_bundle[blockIdx.x];
    const int B = a.bundle_ndir[bundle];
    const int t = threadIdx.x;
    const int b = t % B;
    const int c = t / B;
    const int d = a.bundle_dirs[a.bundle_dir_off[bundle] + b];
    const T cx = a.cxv[d], cy = a.cyv[d];
    const int N = a.N;
    const int4* irec = a.irec + a.item_off[bundle];
    const int4* wc = a.wcell + a.wc_off[2 * bundle];
    const T* lu = a.lu + a.lu_off[bundle];
    T* vdf_d = a.vdf + (size_t)d * N * P;
    const T thermal_on = a.homogeneous ? T(0) : T(1);
    const int l0 = a.lev_off[2 * bundle], l1 = a.lev_off[2 * bundle + 1];

    const int cstep = (int)blockDim.x / B;
    for (int lv = l0; lv < l1; ++lv) {
        for (int s = a.lev_ptr[lv] + c; s < a.lev_ptr[lv + 1]; s += cstep) {
            const int4 r = irec[s];
            const int i = rec_cell(r.x);
            const int fw[3] = {r.y, r.z, r.w};
            const T* cr = a.crec + (size_t)i * a.rs;
            int wk = -1;
#pragma unroll
            for (int il = 0; il < 3; ++il)
                if (code_in(kRecWall, fw_code(fw[il]))) wk = fw_val(fw[il]);
            T x[P];
            if (wk >= 0) {
                const int4 e = wc[wk];
#pragma unroll
                for (int m = 0; m < P; ++m)
                    x[m] = source_plus<P, T>(cr, cx, cy, m, wall_inflow<P, T>(cr, e, cx, cy, a.wtr, thermal_on, m));
            } else {
#pragma unroll
                for (int m = 0; m < P; ++m) x[m] = source_plus<P, T>(cr, cx, cy, m, T(0));
            }
#pragma unroll
            for (int il = 0; il < 3; ++il) {
                const int code = fw_code(fw[il]);
                if (!code_in(kRecNeighbour, code)) continue;
                const T sp = cx * cr[L::R_N + 2 * il] + cy * cr[L::R_N + 2 * il + 1];
                const T w = T(0.5) * (sp - fabs(sp));
                T u[TQ];
                {
                    const bool per = code_in(kRecPeriodic, code);
                    const T* fe = (per && a.use_pbuf)
                                      ? a.pbuf + ((size_t)d * a.F + (bits_of(cr[L::R_INFO + il]) >> 4)) * P
                                      : vdf_d + (size_t)a.other[(size_t)i * 3 + il] * P;
                    const int ilp = fw_ilp(fw[il]);
#pragma unroll
                    for (int q = 0; q < TQ; ++q) u[q] = q < Q ? fe[face_node(DEG, ilp, q < Q ? q : 0)] : T(0);
                }
                face_update<P, T>(x, il, w, a.kref, u);
            }
            const size_t item = (size_t)s * B + b, NB = (size_t)N * B;
            apply_inverse<P, T>(x, [&](int k) { return lu[lu_idx<T>(k, item, NB)]; });
            T* o = vdf_d + (size_t)i * P;
#pragma unroll
            for (int m = 0; m < P; ++m) o[m] = x[m];
        }
        __syncthreads();
    }
}


// the right-hand side of item (slot s, direction b of the bundle) of the
// plain level kernels: source, summed wall inflow, upwind faces, exactly as
// the other sweep kernels assemble it
template <int P, typename T>
__device__ __forceinline__ int otf_rhs(const SweepArgs<T>& a, const int4* irec, const int4* wc, int s, int B, int b,
                                       int d, T cx, T cy, T thermal_on, const T* vdf_d, T (&x)[P])
{
    using L = Lay<P>;
    constexpr int DEG = Deg<P>::v;
    constexpr int Q = DEG + 1;
    constexpr int TQ = Tr<P>::TQ;
    const int4 r = irec[s];
    const int i = rec_cell(r.x);
    const int fw[3] = {r.y, r.z, r.w};
    const T* cr = a.crec + (size_t)i * a.rs;
    int wk = -1;
#pragma unroll
    for (int il = 0; il < 3; ++il)
        if (code_in(kRecWall, fw_code(fw[il]))) wk = fw_val(fw[il]);
    if (wk >= 0) {
        const int4 e = wc[wk];
#pragma unroll
        for (int m = 0; m < P; ++m)
            x[m] = source_plus<P, T>(cr, cx, cy, m, wall_inflow<P, T>(cr, e, cx, cy, a.wtr, thermal_on, m));
    } else {
#pragma unroll
        for (int m = 0; m < P; ++m) x[m] = source_plus<P, T>(cr, cx, cy, m, T(0));
    }
#pragma unroll
    for (int il = 0; il < 3; ++il) {
        const int code = fw_code(fw[il]);
        if (!code_in(kRecNeighbour, code)) continue;
        const T sp = cx * cr[L::R_N + 2 * il] + cy * cr[L::R_N + 2 * il + 1];
        const T w = T(0.5) * (sp - fabs(sp));
        const bool per = code_in(kRecPeriodic, code);
        const T* fe = (per && a.use_pbuf)
                          ? a.pbuf + ((size_t)d * a.F + (bits_of(cr[L::R_INFO + il]) >> 4)) * P
                          : vdf_d + (size_t)a.other[(size_t)i * 3 + il] * P;
        const int ilp = fw_ilp(fw[il]);
        T u[TQ];
#pragma unroll
        for (int q = 0; q < TQ; ++q) u[q] = q < Q ? fe[face_node(DEG, ilp, q < Q ? q : 0)] : T(0);
        face_update<P, T>(x, il, w, a.kref, u);
    }
    return i;
}

// the binary64 sweep of the "mixed" factor mode (binary32 inverses stored,
// binary64 ones not): per item, x = inv32 b refined twice by
// x += inv32 (b - A x), with A x formed in binary64 without A: A is
// area2 (inv_tau_c M - (cx ax + cy ay) Dx - (cx bx + cy by) Dy)
// + sum_faces w len R_face over the reference-element matrices in LDS
// (_kernels.py:106-125 assembles the same operator from the element
// tensors, which are these products).  The inverse's binary32 error
// (about cond(A) 6e-8) shrinks by that factor per refinement: after two,
// x is A^-1 b to binary64 rounding, with no factorisation and no P x P
// matrix in registers
template <int P>
__global__ void __launch_bounds__(256) sweep_refine_kernel(SweepArgs<double> a)
{
    using T = double;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    double* ref = reinterpret_cast<double*>(lds_raw);            // [6][P][P]
    const int bundle = a.block_bundle[blockIdx.x];
    const int B = a.bundle_ndir[bundle];
    const int t = threadIdx.x;
    const int b = t % B;
    const int c = t / B;
    const int d = a.bundle_dirs[a.bundle_dir_off[bundle] + b];
    const T cx = a.cxv[d], cy = a.cyv[d];
    const int N = a.N;
    const int4* irec = a.irec + a.item_off[bundle];
    const int4* wc = a.wcell + a.wc_off[2 * bundle];
    T* vdf_d = a.vdf + (size_t)d * N * P;
    const T thermal_on = a.homogeneous ? T(0) : T(1);
    const int l0 = a.lev_off[2 * bundle], l1 = a.lev_off[2 * bundle + 1];
    const float* inv = a.lu32 + a.lu_off[bundle];
    const size_t NB = (size_t)N * B;
    for (int k = t; k < 6 * P * P; k += blockDim.x) ref[k] = a.refop[k];
    __syncthreads();

    const int cstep = (int)blockDim.x / B;
    for (int lv = l0; lv < l1; ++lv) {
        for (int s = a.lev_ptr[lv] + c; s < a.lev_ptr[lv + 1]; s += cstep) {
            T rhs[P];
            const int i = otf_rhs<P, T>(a, irec, wc, s, B, b, d, cx, cy, thermal_on, vdf_d, rhs);
            const size_t item = (size_t)s * B + b;
            float iv[P * P];
#pragma unroll
            for (int k = 0; k < P * P; ++k) iv[k] = inv[lu_idx<float>(k, item, NB)];
            const double* co = a.cellop + (size_t)i * 8;
            const CellGeo<double> g = a.geo[i];
            double sk[6];
            sk[0] = co[0] * a.inv_tau_c;
            sk[1] = -co[0] * (cx * co[1] + cy * co[3]);
            sk[2] = -co[0] * (cx * co[2] + cy * co[4]);
#pragma unroll
            for (int il = 0; il < 3; ++il) {
                const double sp = cx * g.n[2 * il] + cy * g.n[2 * il + 1];
                sk[3 + il] = 0.5 * (sp + fabs(sp)) * co[5 + il];
            }
            double x[P];
#pragma unroll
            for (int r = 0; r < P; ++r) {
                double v = 0.0;
#pragma unroll
                for (int q = 0; q < P; ++q) v += (double)iv[r * P + q] * rhs[q];
                x[r] = v;
            }
#pragma unroll 1
            for (int it = 0; it < 2; ++it) {
                double res[P];
#pragma unroll
                for (int m = 0; m < P; ++m) {
                    double ax = 0.0;
#pragma unroll
                    for (int l = 0; l < P; ++l) {
                        const int e = m * P + l;
                        const double am = sk[0] * ref[e] + sk[1] * ref[P * P + e] + sk[2] * ref[2 * P * P + e] +
                                          sk[3] * ref[3 * P * P + e] + sk[4] * ref[4 * P * P + e] +
                                          sk[5] * ref[5 * P * P + e];
                        ax += am * x[l];
                    }
                    res[m] = rhs[m] - ax;
                }
#pragma unroll
                for (int r = 0; r < P; ++r) {
                    double v = 0.0;
#pragma unroll
                    for (int q = 0; q < P; ++q) v += (double)iv[r * P + q] * res[q];
                    x[r] += v;
                }
            }
            T* o = vdf_d + (size_t)i * P;
#pragma unroll
            for (int m = 0; m < P; ++m) o[m] = x[m];
        }
        __syncthreads();
    }
}

// the on-the-fly sweep (_kernels.py:242-269): the plain level schedule
// with every element operator built and factorised in binary64 where it is
// used (element_lu), so no factors are stored; the inflow right-hand side
// is assembled exactly as in the other sweep kernels, the solve follows
// DGETRS's loop order
template <int P, typename T>
__global__ void __launch_bounds__(256) sweep_otf_kernel(SweepArgs<T> a)
{
    const int bundle = a.block_bundle[blockIdx.x];
    const int B = a.bundle_ndir[bundle];
    const int t = threadIdx.x;
    const int b = t % B;
    const int c = t / B;
    const int d = a.bundle_dirs[a.bundle_dir_off[bundle] + b];
    const T cx = a.cxv[d], cy = a.cyv[d];
    const double cx64 = a.cxv64[d], cy64 = a.cyv64[d];
    const int N = a.N;
    const int4* irec = a.irec + a.item_off[bundle];
    const int4* wc = a.wcell + a.wc_off[2 * bundle];
    T* vdf_d = a.vdf + (size_t)d * N * P;
    const T thermal_on = a.homogeneous ? T(0) : T(1);
    const int l0 = a.lev_off[2 * bundle], l1 = a.lev_off[2 * bundle + 1];

    const int cstep = (int)blockDim.x / B;
    for (int lv = l0; lv < l1; ++lv) {
        for (int s = a.lev_ptr[lv] + c; s < a.lev_ptr[lv + 1]; s += cstep) {
            T x[P];
            const int i = otf_rhs<P, T>(a, irec, wc, s, B, b, d, cx, cy, thermal_on, vdf_d, x);
            double A[P][P];
            uint64_t pv;
            (void)element_lu<P>(a.mass, a.gx, a.gy, a.fcm, a.geo[i], a.inv_tau_c, N, i, cx64, cy64, A, pv);
            double v[P];
#pragma unroll
            for (int m = 0; m < P; ++m) v[m] = (double)x[m];
            element_lu_solve<P>(A, pv, v);
            T* o = vdf_d + (size_t)i * P;
#pragma unroll
            for (int m = 0; m < P; ++m) o[m] = (T)v[m];
        }
        __syncthreads();
    }
}


// --------------------------------------------------------------------------
// software-pipelined level sweep.  Same per-item arithmetic as sweep_kernel
// (the values agree bit for bit); the instruction stream per level is kept
// short and free of divergent branches on the common path:
//   * records {cell, face words} of level lv+2 and the cell record and
//     inverse of level lv+1 are loaded while level lv is solved (two named
//     register sets, no copies);
//   * an element with downwind readers writes its values on its three
//     faces, face-major, into an LDS pool slot held until its last reader's
//     level has passed; an inflow face is one vector LDS read;
//   * the bundle's inflow wall traces sit in LDS from the start; a face
//     without inflow reads an all-zero slot, so every face runs the same
//     code; only the rare global reads (previous-sweep value of a broken
//     cycle, explicit periodic buffer, overflow bundles) branch.
// --------------------------------------------------------------------------
template <typename T> struct Vec16;
template <> struct Vec16<float> { using type = float4; };
template <> struct Vec16<double> { using type = double2; };

// n contiguous values (n a multiple of 16 / sizeof(T), 16-byte aligned)
// with 16-byte loads
template <typename T, int n>
__device__ __forceinline__ void load16(T* dst, const T* src)
{
    using V = typename Vec16<T>::type;
    constexpr int w = 16 / sizeof(T);
    const V* s = reinterpret_cast<const V*>(src);
#pragma unroll
    for (int k = 0; k < n / w; ++k) {
        const V v = s[k];
        const T* e = reinterpret_cast<const T*>(&v);
#pragma unroll
        for (int j = 0; j < w; ++j) dst[k * w + j] = e[j];
    }
}

template <typename T, int n>
__device__ __forceinline__ void lds_load(T* dst, const T* src)
{
    if constexpr (n * sizeof(T) == 8) {
        using V = typename std::conditional<sizeof(T) == 4, float2, double>::type;
        const V v = *reinterpret_cast<const V*>(src);
        const T* e = reinterpret_cast<const T*>(&v);
#pragma unroll
        for (int j = 0; j < n; ++j) dst[j] = e[j];
    } else {
        load16<T, n>(dst, src);
    }
}

template <typename T, int n>
__device__ __forceinline__ void lds_store(T* dst, const T* src)
{
    if constexpr (n * sizeof(T) == 8) {
        using V = typename std::conditional<sizeof(T) == 4, float2, double>::type;
        V v;
        T* e = reinterpret_cast<T*>(&v);
#pragma unroll
        for (int j = 0; j < n; ++j) e[j] = src[j];
        *reinterpret_cast<V*>(dst) = v;
    } else {
        using V = typename Vec16<T>::type;
        constexpr int w = 16 / sizeof(T);
#pragma unroll
        for (int k = 0; k < n / w; ++k) {
            V v;
            T* e = reinterpret_cast<T*>(&v);
#pragma unroll
            for (int j = 0; j < w; ++j) e[j] = src[k * w + j];
            reinterpret_cast<V*>(dst)[k] = v;
        }
    }
}

struct PipeRec {
    bool valid;
    int slot;
    int4 r;                 // {cell | pool slot, face words}
};

// how the pipelined sweep applies the element operator: the stored inverse
// read in the solve (LU_READ) or gathered with the cell record
// (LU_PREFETCH), or the operator built from the reference-element matrices
// and factorised in registers (LU_FUSED, binary32)
enum { LU_READ = 0, LU_PREFETCH = 1, LU_FUSED = 2, LU_SPLIT = 3 };
template <int P, typename T> constexpr int lu_stored_mode() { return P * P * sizeof(T) <= 288 ? LU_PREFETCH : LU_READ; }

template <int P, typename T, int LUM> struct PipeData {
    using L = Lay<P>;
    bool valid;
    int cell, slot, fw[3];
    int lds[3];         // LDS offset of each face's upwind trace (the zero slot when none)
    int wv;             // LDS offset of the element's wall inflow (the zero slot when none)
    int wst;            // LDS pool slot this item writes, or -1
    T cr[LUM == LU_FUSED ? L::CRF : L::CRL];    // the cell record (without the info words unless fused)
    T lu[LUM == LU_PREFETCH ? L::LU : 4];
};

// what the solve of a level reads from LDS (and, rarely, global memory)
template <int P, typename T> struct PipeIn {
    int code[3];
    T u[3][Tr<P>::TQ];
    T v[pad4(P)];
};

// cache policy of the loads of data read once per sweep (item records,
// element inverses): bit 1 of the gfx940-family buffer-load policy, "nt"
// (non-temporal), so these streams do not push the cell records, which
// every direction reads again, out of L2
constexpr int kStream = 2;

// buffer descriptors of one bundle's streams (raw, bounds-checked: an
// offset past num_records reads 0)
__device__ __forceinline__ __amdgpu_buffer_rsrc_t bundle_rsrc(const void* base, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, bytes, 0x00020000);
}
// 16 loaded bytes (the builtin's native 4 x u32 vector) as another 16-byte type
template <typename R, typename V> __device__ __forceinline__ R bits16(const V& v)
{
    static_assert(sizeof(R) == 16 && sizeof(V) == 16, "16-byte words");
    R r;
    __builtin_memcpy(&r, &v, 16);
    return r;
}

// records of one level: one coalesced 16-byte word per item.  Idle
// threads read the level's first slot (slot 0 for an empty span): every
// load of every stage is unconditional and always consumed by the next
// stage, so the waits the compiler places are exact (a load issued on one
// path only, or left pending on one path, makes a later use or overwrite of
// its registers wait for every prefetch in flight)
__device__ __forceinline__ void pipe_record(__amdgpu_buffer_rsrc_t irec, int2 span, int c, int B, int b, PipeRec& o)
{
    const int sl = span.x + c;
    o.valid = sl < span.y;
    o.slot = o.valid ? sl : span.x;
    o.r = bits16<int4>(__builtin_amdgcn_raw_buffer_load_b128(irec, (uint32_t)o.slot * 16u, 0, kStream));
}

// record-code classes as bit masks over the code: one shift and one test,
// no short-circuit control flow
// faces whose inflow trace comes from the LDS pool
// item (slot, direction b) of a bundle of B directions as an LDS row index:
// slot B + b, in 24-bit multiplies (slots and wall-list indices are far
// below 2^24) so that no 64-bit multiply-add is needed; one direction per
// bundle: the slot itself
template <int BF> __device__ __forceinline__ unsigned bundle_row(unsigned slot, int B, int b)
{
    if constexpr (BF) return slot;
    else return __umul24(slot, (unsigned)B) + (unsigned)b;
}

// how a face word selects its LDS trace (masks over the record codes,
// uniform per launch): the codes read from LDS, those whose slot is in bits
// 13-25, those whose slot counts from the first stage slot
struct FaceSel { int lds, hi, stage; };
__device__ __forceinline__ FaceSel face_sel(int use_pbuf, int stage_on)
{
    FaceSel f{code_bit(REC_UP_LDS), 0, 0};
    if (!use_pbuf) f.lds |= code_bit(REC_PER_LDS) | code_bit(REC_PER_LDS_ST);
    if (stage_on) {
        f.lds |= kRecStaged;
        f.stage = code_bit(REC_OLD_ST) | code_bit(REC_PER_OLD_ST) | (use_pbuf ? code_bit(REC_PER_LDS_ST) : 0);
        f.hi = use_pbuf ? code_bit(REC_PER_OLD_ST) | code_bit(REC_PER_LDS_ST) : 0;
    }
    return f;
}

// the prologue's copy of the staged traces: per stage slot and direction of
// the bundle the one face trace a staged read takes, from the previous
// sweep's values of the element across or, for a periodic face with an
// explicit periodic buffer, from the buffer's entry of the face.  A class's
// list holds its staged reads in three runs: previous-sweep values across
// non-periodic faces, across periodic faces, then every periodic face (the
// buffer's entries); the stage slots are the first two runs without an
// explicit periodic buffer, the first and the third with one.
template <int P, typename T>
__device__ __forceinline__ void stage_traces(const SweepArgs<T>& a, int bundle, int B, int t, int nthr, T* lds)
{
    constexpr int DEG = Deg<P>::v;
    constexpr int Q = DEG + 1;
    constexpr int TQ = Tr<P>::TQ;
    constexpr int W = Tr<P>::W;
    if (!a.stage_on) return;
    const int4 so = a.stage_off[bundle];
    const int n = so.y + (a.use_pbuf ? so.w : so.z);
    const int32_t* lst = a.stage_list + so.x;
    T* st = lds + (size_t)(a.pool + 2) * B * W;
    const int32_t* dirs = a.bundle_dirs + a.bundle_dir_off[bundle];
    for (int k = t; k < n * B * TQ; k += nthr) {
        const int e = k / (B * TQ), bb = (k / TQ) % B, q = k % TQ;
        const bool buf = a.use_pbuf && e >= so.y;
        const int ent = lst[buf ? e + so.z : e];
        const int idx = ent >> 2, ilp = ent & 3;
        const int db = dirs[bb];
        const T* src = buf ? a.pbuf + ((size_t)db * a.F + idx) * P : a.vdf + ((size_t)db * a.N + idx) * P;
        st[k] = q < Q ? src[face_node(DEG, ilp, q < Q ? q : 0)] : T(0);
    }
}

// gather: only loads whose addresses come from the record of two levels
// earlier; every use of the loaded data waits for the solve stage (the
// in-order vmcnt counter would make an earlier use wait for every
// prefetch in flight).  The inverse streams through a buffer descriptor:
// one lane offset, the k-groups a scalar offset apart.
template <int P, typename T, int LUM, int BF>
__device__ __forceinline__ void pipe_gather(const SweepArgs<T>& a, const PipeRec& r, int B, int b,
                                            __amdgpu_buffer_rsrc_t lu, PipeData<P, T, LUM>& o)
{
    using L = Lay<P>;
    constexpr int W = Tr<P>::W;
    constexpr int TQ = Tr<P>::TQ;
    constexpr int PV = pad4(P);
    o.valid = r.valid;
    const int i = rec_cell(r.r.x);
    o.cell = i;
    o.slot = r.slot;
    o.wst = rec_pool(r.r.x);
    o.fw[0] = r.r.y;
    o.fw[1] = r.r.z;
    o.fw[2] = r.r.w;
    const unsigned zero = ((unsigned)a.pool * B + b) * W;
    const FaceSel fs = face_sel(a.use_pbuf, a.stage_on);
    const unsigned stage0 = ((unsigned)a.pool + 2u) * B * W + (unsigned)b * TQ;
    unsigned wv = zero;
#pragma unroll
    for (int il = 0; il < 3; ++il) {
        // a pool or stage face reads its slot, every other face the all-zero
        // slot; a wall face names the element's wall inflow (selects, no
        // control flow)
        const int code = fw_code(o.fw[il]);
        const unsigned val = (unsigned)fw_val(o.fw[il]);
        const unsigned slot = code_in(fs.hi, code) ? val >> kStageBits : val & (unsigned)kStageMask;
        const bool st = code_in(fs.stage, code);
        unsigned nbr;
        if constexpr (BF) {
            // one direction per bundle: a stage slot holds one trace, a pool
            // slot three; one multiply-add with selected constants
            nbr = slot * (st ? (unsigned)TQ : (unsigned)W) + (st ? stage0 : (unsigned)fw_ilp(o.fw[il]) * TQ);
        } else {
            // (the selected-factor form measured slower in bundles)
            nbr = st ? stage0 + __umul24(slot, (unsigned)(B * TQ))
                     : __umul24(bundle_row<BF>(slot, B, b), (unsigned)W) + (unsigned)fw_ilp(o.fw[il]) * TQ;
        }
        o.lds[il] = (int)(code_in(fs.lds, code) ? nbr : zero);
        // a wall face's value is its element's index in the bundle's wall
        // list, far below 2^24 (the list lives in LDS)
        wv = code_in(kRecWall, code) ? (unsigned)a.wbase + bundle_row<BF>(val & 0xffffffu, B, b) * PV : wv;
    }
    o.wv = (int)wv;
    // the element's row: element below 2^21, stride below 2^8
    load16<T, (LUM == LU_FUSED ? L::CRF : L::CRL)>(o.cr, a.crec + __umul24((unsigned)i, (unsigned)a.rs));
    if constexpr (LUM == LU_PREFETCH) {
        constexpr int w = 16 / sizeof(T);
        const uint32_t voff = (uint32_t)(r.slot * B + b) * 16u;
        const uint32_t gstride = (uint32_t)a.N * (uint32_t)B * 16u;
#pragma unroll
        for (int g = 0; g < L::LU / w; ++g) {
            const auto v = bits16<typename Vec16<T>::type>(
                __builtin_amdgcn_raw_buffer_load_b128(lu, voff, (uint32_t)g * gstride, kStream));
            const T* t = reinterpret_cast<const T*>(&v);
#pragma unroll
            for (int j = 0; j < w; ++j) o.lu[g * w + j] = t[j];
        }
    }
}

// first half of the solve: the upwind traces and the wall inflow from LDS
// (issued before the next gather, so their latency passes while it issues)
// and, in the kernel instance launched when such reads exist, the rare
// global reads (explicit periodic buffer, a previous-sweep value of a
// broken cycle, a bundle running with full barriers) behind a wave-uniform
// branch with the wait inside it
template <int P, typename T, int LUM, bool GREAD>
__device__ __forceinline__ void pipe_read(const SweepArgs<T>& a, const PipeData<P, T, LUM>& dC, int d, const T* lds,
                                          const T* vdf_d, PipeIn<P, T>& in)
{
    using L = Lay<P>;
    constexpr int DEG = Deg<P>::v;
    constexpr int Q = DEG + 1;
    constexpr int TQ = Tr<P>::TQ;
    constexpr int PV = pad4(P);
#pragma unroll
    for (int il = 0; il < 3; ++il) {
        in.code[il] = fw_code(dC.fw[il]);
        lds_load<T, TQ>(in.u[il], lds + dC.lds[il]);
    }
    lds_load<T, PV>(in.v, lds + dC.wv);
    if constexpr (GREAD) {
        const int gmask = dC.valid ? kRecNeighbour & ~face_sel(a.use_pbuf, a.stage_on).lds : 0;
        const bool g0 = code_in(gmask, in.code[0]), g1 = code_in(gmask, in.code[1]),
                   g2 = code_in(gmask, in.code[2]);
        if (__builtin_amdgcn_ballot_w64(g0 || g1 || g2) != 0) {
#pragma unroll
            for (int il = 0; il < 3; ++il) {
                if (code_in(gmask, in.code[il])) {
                    const int ilp = fw_ilp(dC.fw[il]);
                    // a staged face's value names a stage slot: the
                    // element across comes from the neighbour list
                    const int nb = code_in(kRecStaged, in.code[il]) ? a.other[(size_t)dC.cell * 3 + il]
                                                                    : fw_val(dC.fw[il]);
                    const T* fe = (code_in(kRecPeriodic, in.code[il]) && a.use_pbuf)
                                      ? a.pbuf + ((size_t)d * a.F +
                                                  (bits_of(a.crec[(size_t)dC.cell * a.rs + L::R_INFO + il]) >> 4)) * P
                                      : vdf_d + (size_t)nb * P;
#pragma unroll
                    for (int q = 0; q < Q; ++q) in.u[il][q] = fe[face_node(DEG, ilp, q)];
                }
            }
            __builtin_amdgcn_s_waitcnt(0x0f70);       // vmcnt(0)
        }
    }
}

// second half: source, wall inflow, upwind faces, the inverse, the stores
template <int P, typename T, int LUM, int BF>
__device__ __forceinline__ void pipe_compute(const SweepArgs<T>& a, const PipeData<P, T, LUM>& dC,
                                             const PipeIn<P, T>& in, T cx, T cy, int B, int b, T* lds, T* vdf_d,
                                             const T* lub, const T* fv = nullptr)
{
    using L = Lay<P>;
    constexpr int DEG = Deg<P>::v;
    constexpr int Q = DEG + 1;
    constexpr int TQ = Tr<P>::TQ;
    const T* cr = dC.cr;
    // idle threads of a busy wave compute on the first slot's data and
    // store nothing that is read
    T x[P];
#pragma unroll
    for (int m = 0; m < P; ++m) x[m] = source_plus<P, T>(cr, cx, cy, m, in.v[m]);
#pragma unroll
    for (int il = 0; il < 3; ++il) {
        const T sp = cx * cr[L::R_N + 2 * il] + cy * cr[L::R_N + 2 * il + 1];
        const T w = T(0.5) * (sp - fabs(sp));
        face_update<P, T>(x, il, w, a.kref, in.u[il]);
    }
    if constexpr (LUM == LU_FUSED) {
        fused_solve<P, T>(x, cr, cx, cy);
    } else if constexpr (LUM == LU_SPLIT) {
        fused_substitute<P, T>(x, [&](int i, int j) { return fv[i * P + j]; });
    } else if constexpr (LUM == LU_PREFETCH) {
        apply_inverse<P, T>(x, [&](int k) { return dC.lu[k]; });
    } else {
        const size_t item = (size_t)dC.slot * B + b, NB = (size_t)a.N * B;
        apply_inverse<P, T>(x, [&](int k) { return __builtin_nontemporal_load(lub + lu_idx<T>(k, item, NB)); });
    }
    // stores without a branch: an idle thread, and an element nobody reads
    // through LDS, write their traces to the spill slot; an idle thread
    // writes its values to a scratch row (a branch would let the compiler
    // sink the whole solve into it, leaving the loads of idle threads
    // unconsumed)
    {
        constexpr int W = Tr<P>::W;
        const int spill = a.pool + 1;
        const int sl = (dC.valid && dC.wst >= 0) ? dC.wst : spill;
        const int wr = (int)(bundle_row<BF>((unsigned)sl, B, b) * W);
        T* o = dC.valid ? vdf_d + (size_t)dC.cell * P : a.spill_row;
#pragma unroll
        for (int m = 0; m < P; ++m) o[m] = x[m];
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            T t[TQ];
#pragma unroll
            for (int q = 0; q < TQ; ++q) t[q] = q < Q ? x[face_node(DEG, f, q < Q ? q : 0)] : T(0);
            lds_store<T, TQ>(lds + wr + f * TQ, t);
        }
    }
}

__device__ __forceinline__ void level_barrier(bool global_mode)
{
    if (global_mode) {
        __syncthreads();
    } else {
        // every value exchanged within the sweep lives in LDS: order LDS
        // traffic only and leave the prefetch loads in flight
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    }
}

// threads of a pipelined workgroup at most (one item each per level)
template <typename T> struct PipeThreads {
    static constexpr int max = 256;
};

// BF: 1 when every bundle holds one direction (the default schedule; the
// slot and LDS arithmetic then drops its bundle factors), 0 for bundles
template <int P, typename T, int LUM, bool GREAD, int DEPTH, int BF>
__device__ __forceinline__ void sweep_pipe_body(const SweepArgs<T>& a, int blk)
{
    constexpr int W = Tr<P>::W;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int bundle = a.block_bundle[blk];
    const int B = BF ? 1 : a.bundle_ndir[bundle];
    // LDS: pool slots, one all-zero slot, one spill slot, the stage slots, the wall inflow
    T* lds = reinterpret_cast<T*>(lds_raw);
    const bool global_mode = a.bundle_global[bundle] != 0;
    const int t = threadIdx.x;
    const int b = BF ? 0 : t % B;
    const int c = BF ? t : t / B;
    const int d = a.bundle_dirs[a.bundle_dir_off[bundle] + b];
    const T cx = a.cxv[d], cy = a.cyv[d];
    const int N = a.N;
    const int64_t it0 = a.item_off[bundle];
    const __amdgpu_buffer_rsrc_t irec = bundle_rsrc(a.irec + it0, (uint32_t)N * 16u);
    const int64_t lu0 = a.lu_off[bundle];
    const __amdgpu_buffer_rsrc_t lu = bundle_rsrc(a.lu + lu0, (uint32_t)((a.lu_off[bundle + 1] - lu0) * sizeof(T)));
    T* vdf_d = a.vdf + (size_t)d * N * P;
    const int l0 = a.lev_off[2 * bundle], l1 = a.lev_off[2 * bundle + 1];
    // waves past the context's own workgroup (a batch launches the widest
    // one) sweep nothing: they take the prologue barrier and then as many
    // level barriers as the sweeping waves.  nthr is a multiple of 64, and
    // the test reads lane 0's index, so the branch is uniform per wave and
    // scalar (no barrier of the other path runs under an empty exec mask).
    const int nthr = a.nthr;
    const bool sweeps = __builtin_amdgcn_readfirstlane(t) < nthr;
    if (sweeps) {
        T* z = lds + (size_t)a.pool * B * W;
        for (int k = t; k < B * W; k += nthr) z[k] = T(0);
        // the summed wall inflow of every inflow wall element of the bundle,
        // per direction (a thermal wall emits nothing in a homogeneous sweep)
        constexpr int PV = pad4(P);
        const int w0 = a.wc_off[2 * bundle], nw = a.wc_off[2 * bundle + 1];
        const T thermal_on = a.homogeneous ? T(0) : T(1);
        for (int k = t; k < nw * B * PV; k += nthr) {
            const int e = k / (B * PV), bb = (k / PV) % B, m = k % PV;
            const int4 ent = a.wcell[w0 + e];
            const int db = a.bundle_dirs[a.bundle_dir_off[bundle] + bb];
            const T* cr = a.crec + (size_t)ent.x * a.rs;
            lds[a.wbase + k] = m < P ? wall_inflow<P, T>(cr, ent, a.cxv[db], a.cyv[db], a.wtr, thermal_on, m) : T(0);
        }
        stage_traces<P, T>(a, bundle, B, t, nthr, lds);
    }
    __syncthreads();
    if (!sweeps) {
        // the loops below run whole trips (two levels at depth 2, six at
        // depth 3), one barrier per level
        const int nlev = l1 > l0 ? l1 - l0 : 0;
        const int trip = DEPTH == 4 ? 12 : (DEPTH == 3 ? 6 : DEPTH);
        const int nbar = (nlev + trip - 1) / trip * trip;
        for (int k = 0; k < nbar; ++k) level_barrier(global_mode);
        return;
    }

    // The spans are read with per-lane (vector) loads: a scalar load would
    // share lgkmcnt with the LDS traffic that every level waits for.  A
    // level past the end reads the last level's span and takes (0, 0): the
    // load is unconditional (a load on one path only leaves the compiler
    // unsure of the in-order vmcnt and it then waits for every prefetch in
    // flight).
    const int32_t* lp = a.lev_ptr + (t & a.lane_zero);
    auto span = [&](int lv) {
        const int k = min(lv, l1 - 1);
        const int2 v = make_int2(lp[k], lp[k + 1]);
        return lv < l1 ? v : make_int2(0, 0);
    };
    // one level: in binary32 the solve's LDS reads, the loads of later
    // levels, then the solve's arithmetic and stores (the LDS latency passes
    // while the loads issue); in binary64 the loads first (the read values
    // held across the gather would push the binary64 register sets into
    // the accumulation registers)
    auto level = [&](const PipeData<P, T, LUM>& dS, PipeRec& rN, int2& sN, int lvN, PipeData<P, T, LUM>& dG,
                     const PipeRec& rG) {
        PipeIn<P, T> in;
        if constexpr (sizeof(T) == 4) pipe_read<P, T, LUM, GREAD>(a, dS, d, lds, vdf_d, in);
        pipe_record(irec, sN, c, B, b, rN);
        sN = span(lvN);
        pipe_gather<P, T, LUM, BF>(a, rG, B, b, lu, dG);
        if constexpr (sizeof(T) != 4) pipe_read<P, T, LUM, GREAD>(a, dS, d, lds, vdf_d, in);
        pipe_compute<P, T, LUM, BF>(a, dS, in, cx, cy, B, b, lds, vdf_d, a.lu + lu0);
        level_barrier(global_mode);
    };
    if constexpr (DEPTH == 2) {
        // level spans travel two levels ahead of their use, the records one
        // level ahead of the gather, the gather one level ahead of the
        // solve; every stage alternates between two named register sets so
        // that no copy forces a wait
        int2 S0 = span(l0 + 2), S1 = span(l0 + 3);
        PipeRec R0, R1;
        PipeData<P, T, LUM> D0, D1;
        // the prologue issues its loads in the steady-state order (records
        // of the level after next before the gather of the next level):
        // vmcnt is in order, and the loop-entry state must not ask for more
        // waiting
        pipe_record(irec, span(l0), c, B, b, R0);
        pipe_record(irec, span(l0 + 1), c, B, b, R1);
        pipe_gather<P, T, LUM, BF>(a, R0, B, b, lu, D0);
        // two levels per trip, with no exit between them: an odd level count
        // runs one empty level (span (0, 0)) at the end.  A mid-trip exit
        // would make the compiler rotate the register sets with copies, and a
        // copy of a loaded value waits for it.
        for (int lv = l0; lv < l1; lv += 2) {
            level(D0, R0, S0, lv + 4, D1, R1);     // solve lv; record lv+2; gather lv+1
            level(D1, R1, S1, lv + 5, D0, R0);     // solve lv+1; record lv+3; gather lv+2
        }
    } else if constexpr (DEPTH == 4) {
        // the gather three levels ahead of the solve and the records two
        // levels ahead of the gather (four data sets, three record and span
        // sets, twelve levels per trip): every gathered load has three level
        // times to arrive.  Data set k holds level l0 + k mod 4, record and
        // span set k the levels l0 + k mod 3.
        int2 S0 = span(l0 + 6), S1 = span(l0 + 7), S2 = span(l0 + 5);
        PipeRec R0, R1, R2;
        PipeData<P, T, LUM> D0, D1, D2, D3;
        pipe_record(irec, span(l0), c, B, b, R0);
        pipe_record(irec, span(l0 + 1), c, B, b, R1);
        pipe_gather<P, T, LUM, BF>(a, R0, B, b, lu, D0);
        pipe_record(irec, span(l0 + 2), c, B, b, R2);
        pipe_gather<P, T, LUM, BF>(a, R1, B, b, lu, D1);
        pipe_record(irec, span(l0 + 3), c, B, b, R0);
        pipe_gather<P, T, LUM, BF>(a, R2, B, b, lu, D2);
        pipe_record(irec, span(l0 + 4), c, B, b, R1);
        for (int lv = l0; lv < l1; lv += 12) {
            level(D0, R2, S2, lv + 8, D3, R0);     // solve lv; record lv+5; gather lv+3
            level(D1, R0, S0, lv + 9, D0, R1);     // solve lv+1; record lv+6; gather lv+4
            level(D2, R1, S1, lv + 10, D1, R2);
            level(D3, R2, S2, lv + 11, D2, R0);
            level(D0, R0, S0, lv + 12, D3, R1);
            level(D1, R1, S1, lv + 13, D0, R2);
            level(D2, R2, S2, lv + 14, D1, R0);
            level(D3, R0, S0, lv + 15, D2, R1);
            level(D0, R1, S1, lv + 16, D3, R2);
            level(D1, R2, S2, lv + 17, D0, R0);
            level(D2, R0, S0, lv + 18, D1, R1);
            level(D3, R1, S1, lv + 19, D2, R2);
        }
    } else {
        // the gather two levels ahead of the solve and the records two levels
        // ahead of the gather (three register sets each, six levels per
        // trip: the loop's back edge, where the sets are renamed, costs a
        // wait for the loads in flight once per trip): every load has two
        // level times to arrive.  Set k holds level l0 + k mod 3.
        int2 S0 = span(l0 + 6), S1 = span(l0 + 4), S2 = span(l0 + 5);
        PipeRec R0, R1, R2;
        PipeData<P, T, LUM> D0, D1, D2;
        pipe_record(irec, span(l0), c, B, b, R0);
        pipe_record(irec, span(l0 + 1), c, B, b, R1);
        pipe_gather<P, T, LUM, BF>(a, R0, B, b, lu, D0);
        pipe_record(irec, span(l0 + 2), c, B, b, R2);
        pipe_gather<P, T, LUM, BF>(a, R1, B, b, lu, D1);
        pipe_record(irec, span(l0 + 3), c, B, b, R0);
        for (int lv = l0; lv < l1; lv += 6) {
            level(D0, R1, S1, lv + 7, D2, R2);     // solve lv; record lv+4; gather lv+2
            level(D1, R2, S2, lv + 8, D0, R0);     // solve lv+1; record lv+5; gather lv+3
            level(D2, R0, S0, lv + 9, D1, R1);     // solve lv+2; record lv+6; gather lv+4
            level(D0, R1, S1, lv + 10, D2, R2);
            level(D1, R2, S2, lv + 11, D0, R0);
            level(D2, R0, S0, lv + 12, D1, R1);
        }
    }
}

// ----------------------------------------------------------------------------
// the split fused sweep (binary32, P <= 6): a workgroup of twice the level
// width, its first half of waves solving level lv (upwind traces, right-hand
// side, substitution, stores) while its second half builds and factorises
// the element operators of level lv + 1 and leaves them in one of two LDS
// factor buffers, [buffer][group of 4 values][lane][4].  The consumer's
// data (records, the cell record without the element scalars) and the
// producer's (records, the normals and element scalars) each stream in
// their own three-level pipeline; both halves take the same barriers, one
// per level and one for the first level's factors.  The two halves of a
// SIMD interleave: the factorisation runs while the solve waits on memory
// and barriers.  The values are the one-wave fused sweep's, bit for bit
// (the same build, elimination and substitution).
// ----------------------------------------------------------------------------
template <int P> constexpr int fac_vals() { return pad4(P * P); }
template <int P> constexpr int prod_lo() { return (3 * P) & ~3; }                    // first value a producer loads
template <int P> constexpr int prod_vals() { return pad4(3 * P + 10) - prod_lo<P>(); }

template <int P, typename T> struct ProdData {
    T g[prod_vals<P>()];     // cell record values prod_lo .. pad4(3P + 14)
};

template <int P, typename T>
__device__ __forceinline__ void prod_gather(const SweepArgs<T>& a, const PipeRec& r, ProdData<P, T>& o)
{
    load16<T, prod_vals<P>()>(o.g, a.crec + (size_t)rec_cell(r.r.x) * a.rs + prod_lo<P>());
}

template <int P, typename T>
__device__ __forceinline__ void prod_factor(const ProdData<P, T>& dP, T cx, T cy, T* fac, int stride)
{
    using L = Lay<P>;
    constexpr int FV = fac_vals<P>();
    T A[P][P];
    fused_build<P, T>(A, dP.g[L::R_FUSE - prod_lo<P>()], dP.g + (L::R_N - prod_lo<P>()), cx, cy);
    fused_eliminate<P, T>(A);
    T f[FV];
#pragma unroll
    for (int k = 0; k < FV; ++k) f[k] = k < P * P ? A[k / P][k % P] : T(0);
#pragma unroll
    for (int g = 0; g < FV / 4; ++g) lds_store<T, 4>(fac + (size_t)g * stride, f + 4 * g);
}

template <int P, typename T, bool GREAD, int BF>
__device__ __forceinline__ void sweep_split_body(const SweepArgs<T>& a, int blk)
{
    constexpr int FV = fac_vals<P>();
    constexpr int W = Tr<P>::W;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int bundle = a.block_bundle[blk];
    const int B = BF ? 1 : a.bundle_ndir[bundle];
    T* lds = reinterpret_cast<T*>(lds_raw);
    const bool global_mode = a.bundle_global[bundle] != 0;
    const int half = (int)blockDim.x / 2;
    const int t = threadIdx.x;
    // the halves are whole waves (half is a multiple of 64): uniform per wave
    const bool producer = __builtin_amdgcn_readfirstlane(t) >= half;
    const int tr = producer ? t - half : t;
    const int b = BF ? 0 : tr % B;
    const int c = BF ? tr : tr / B;
    const int d = a.bundle_dirs[a.bundle_dir_off[bundle] + b];
    const T cx = a.cxv[d], cy = a.cyv[d];
    const int N = a.N;
    const int64_t it0 = a.item_off[bundle];
    const __amdgpu_buffer_rsrc_t irec = bundle_rsrc(a.irec + it0, (uint32_t)N * 16u);
    T* vdf_d = a.vdf + (size_t)d * N * P;
    const int l0 = a.lev_off[2 * bundle], l1 = a.lev_off[2 * bundle + 1];
    const int nthr = a.nthr;
    const bool sweeps = __builtin_amdgcn_readfirstlane(tr) < nthr;
    if (sweeps && !producer) {
        T* z = lds + (size_t)a.pool * B * W;
        for (int k = tr; k < B * W; k += nthr) z[k] = T(0);
        constexpr int PV = pad4(P);
        const int w0 = a.wc_off[2 * bundle], nw = a.wc_off[2 * bundle + 1];
        const T thermal_on = a.homogeneous ? T(0) : T(1);
        for (int k = tr; k < nw * B * PV; k += nthr) {
            const int e = k / (B * PV), bb = (k / PV) % B, m = k % PV;
            const int4 ent = a.wcell[w0 + e];
            const int db = a.bundle_dirs[a.bundle_dir_off[bundle] + bb];
            const T* cr = a.crec + (size_t)ent.x * a.rs;
            lds[a.wbase + k] = m < P ? wall_inflow<P, T>(cr, ent, a.cxv[db], a.cyv[db], a.wtr, thermal_on, m) : T(0);
        }
        stage_traces<P, T>(a, bundle, B, tr, nthr, lds);
    }
    __syncthreads();
    // the lane's factor values in buffer k: group g at fac(k) + g * 4 * half
    T* fac_base = lds + a.fbase + (size_t)tr * 4;
    const int fstride = 4 * half;
    auto fac = [&](int lv) { return fac_base + (size_t)((lv - l0) & 1) * (FV / 4) * fstride; };
    const int nlev = l1 > l0 ? l1 - l0 : 0;
    const int ntrip = (nlev + 2) / 3;
    if (!sweeps) {
        for (int k = 0; k < 1 + 3 * ntrip; ++k) level_barrier(global_mode);
        return;
    }
    const int32_t* lp = a.lev_ptr + (t & a.lane_zero);
    auto span = [&](int lv) {
        const int k = min(lv, l1 - 1);
        const int2 v = make_int2(lp[k], lp[k + 1]);
        return lv < l1 ? v : make_int2(0, 0);
    };
    if (producer) {
        // production n: the factors of level l0 + n into buffer n & 1 while
        // the consumers solve level l0 + n - 1; its records were read four
        // productions ahead, its cell data gathered two ahead (register set
        // n mod 3); one barrier each, 1 + 3 ntrip in all
        auto produce = [&](const ProdData<P, T>& dS, PipeRec& rN, int2& sN, int lvN, ProdData<P, T>& dG,
                           const PipeRec& rG, int p) {
            pipe_record(irec, sN, c, B, b, rN);
            sN = span(lvN);
            prod_gather<P, T>(a, rG, dG);
            prod_factor<P, T>(dS, cx, cy, fac(p), fstride);
            level_barrier(global_mode);
        };
        int2 S0 = span(l0 + 6), S1 = span(l0 + 4), S2 = span(l0 + 5);
        PipeRec R0, R1, R2;
        ProdData<P, T> D0, D1, D2;
        pipe_record(irec, span(l0), c, B, b, R0);
        pipe_record(irec, span(l0 + 1), c, B, b, R1);
        prod_gather<P, T>(a, R0, D0);
        pipe_record(irec, span(l0 + 2), c, B, b, R2);
        prod_gather<P, T>(a, R1, D1);
        pipe_record(irec, span(l0 + 3), c, B, b, R0);
        produce(D0, R1, S1, l0 + 7, D2, R2, l0);
        for (int n = 1; n < 1 + 3 * ntrip; n += 3) {
            produce(D1, R2, S2, l0 + n + 7, D0, R0, l0 + n);
            produce(D2, R0, S0, l0 + n + 8, D1, R1, l0 + n + 1);
            produce(D0, R1, S1, l0 + n + 9, D2, R2, l0 + n + 2);
        }
    } else {
        auto level = [&](const PipeData<P, T, LU_SPLIT>& dS, PipeRec& rN, int2& sN, int lvN,
                         PipeData<P, T, LU_SPLIT>& dG, const PipeRec& rG, int lv) {
            PipeIn<P, T> in;
            pipe_read<P, T, LU_SPLIT, GREAD>(a, dS, d, lds, vdf_d, in);
            T fv[FV];
            const T* fb = fac(lv);
#pragma unroll
            for (int g = 0; g < FV / 4; ++g) lds_load<T, 4>(fv + 4 * g, fb + (size_t)g * fstride);
            pipe_record(irec, sN, c, B, b, rN);
            sN = span(lvN);
            pipe_gather<P, T, LU_SPLIT, BF>(a, rG, B, b, irec, dG);
            pipe_compute<P, T, LU_SPLIT, BF>(a, dS, in, cx, cy, B, b, lds, vdf_d, nullptr, fv);
            level_barrier(global_mode);
        };
        level_barrier(global_mode);            // the producers' first level
        int2 S0 = span(l0 + 6), S1 = span(l0 + 4), S2 = span(l0 + 5);
        PipeRec R0, R1, R2;
        PipeData<P, T, LU_SPLIT> D0, D1, D2;
        pipe_record(irec, span(l0), c, B, b, R0);
        pipe_record(irec, span(l0 + 1), c, B, b, R1);
        pipe_gather<P, T, LU_SPLIT, BF>(a, R0, B, b, irec, D0);
        pipe_record(irec, span(l0 + 2), c, B, b, R2);
        pipe_gather<P, T, LU_SPLIT, BF>(a, R1, B, b, irec, D1);
        pipe_record(irec, span(l0 + 3), c, B, b, R0);
        for (int n = 0; n < 3 * ntrip; n += 3) {
            level(D0, R1, S1, l0 + n + 7, D2, R2, l0 + n);
            level(D1, R2, S2, l0 + n + 8, D0, R0, l0 + n + 1);
            level(D2, R0, S0, l0 + n + 9, D1, R1, l0 + n + 2);
        }
    }
}

template <int P, typename T, bool GREAD, int BF>
__global__ void __launch_bounds__(512) sweep_split_kernel(SweepArgs<T> a)
{
    sweep_split_body<P, T, GREAD, BF>(a, blockIdx.x);
}

template <int P, typename T, int LUM, bool GREAD, int DEPTH, int BF>
__global__ void __launch_bounds__(PipeThreads<T>::max) sweep_pipe_kernel(SweepArgs<T> a)
{
    sweep_pipe_body<P, T, LUM, GREAD, DEPTH, BF>(a, blockIdx.x);
}

// several contexts' sweeps in one persistent launch fed by a device work
// queue: the work items are the (context, block) pairs of bmap, longest
// bundles first; each resident workgroup takes the next item from the
// queue head (one atomic per item), sweeps block bmap[k].y of context
// bmap[k].x with that context's arguments, and stops once the queue is
// empty.  The head is zeroed by queue_reset_kernel before the launch.
template <int P, typename T, int LUM, bool GREAD, int DEPTH, int BF>
__global__ void __launch_bounds__(PipeThreads<T>::max) sweep_queue_kernel(const SweepArgs<T>* __restrict__ args,
                                                          const int2* __restrict__ bmap, int nitems,
                                                          int* __restrict__ head)
{
    __shared__ int next;
    for (;;) {
        // every wave is past the previous item (its LDS pool and `next`)
        __syncthreads();
        if (threadIdx.x == 0) next = atomicAdd(head, 1);
        __syncthreads();
        const int k = next;
        if (k >= nitems) break;
        const int2 m = bmap[k];
        sweep_pipe_body<P, T, LUM, GREAD, DEPTH, BF>(args[m.x], m.y);
    }
}

__global__ void queue_reset_kernel(int* head) { *head = 0; }

// the wall vector of each adiabatic wall face from this iterate's emission
// flux_wall (bc.py:47-58, _kernels.py:364-397), over the face length
template <int P, typename T>
__global__ void __launch_bounds__(256) wall_vec_kernel(const double* __restrict__ flux_wall,
                                                       const int32_t* __restrict__ row,
                                                       const int32_t* __restrict__ face,
                                                       const double* __restrict__ rlen, int n,
                                                       T* __restrict__ wtr)
{
    constexpr int PV = pad4(P);
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n * P) return;
    const int k = g / P, m = g % P;
    wtr[(size_t)row[k] * PV + m] = (T)(flux_wall[(size_t)face[k] * P + m] * rlen[k]);
}

// ==========================================================================
// moments: per (cell, dof) the direction sums in index order, binary64
// (_kernels.py:311-338); element integrals in a second pass
// ==========================================================================
template <typename T>
__global__ void __launch_bounds__(256) moments_kernel(const T* __restrict__ vdf,
                                                      const double* __restrict__ cxv,
                                                      const double* __restrict__ cyv,
                                                      const double* __restrict__ dw, int D, int N,
                                                      int P, double cv, double* ts, double* qxs,
                                                      double* qys, double* ssraw)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= N * P) return;
    const int i = g / P, l = g - (g / P) * P;
    double ss = 0.0, ssx = 0.0, ssy = 0.0;
    const size_t stride = (size_t)N * P;
    const T* f = vdf + g;
    for (int d = 0; d < D; ++d) {
        const double v = (double)f[(size_t)d * stride];
        const double w = dw[d];
        ss += v * w;
        ssx += cxv[d] * v * w;
        ssy += cyv[d] * v * w;
    }
    ts[(size_t)l * N + i] = ss / cv;
    qxs[(size_t)l * N + i] = ssx;
    qys[(size_t)l * N + i] = ssy;
    ssraw[(size_t)l * N + i] = ss;
}

// the same sums, V consecutive (cell, dof) values per thread read as one
// 16-byte word per direction, the words of 16 (binary32) or 8 (binary64)
// directions in flight before the in-order accumulation (each value's sum
// keeps the order and expressions of moments_kernel).  Needs N*P % V == 0
// and a 16-byte aligned vdf.
template <typename T, int V>
__global__ void __launch_bounds__(256) moments_vec_kernel(const T* __restrict__ vdf,
                                                          const double* __restrict__ cxv,
                                                          const double* __restrict__ cyv,
                                                          const double* __restrict__ dw, int D, int N,
                                                          int P, double cv, double* ts, double* qxs,
                                                          double* qys, double* ssraw)
{
    static_assert(V * sizeof(T) == 16, "one 16-byte word per direction");
    using W = typename Vec16<T>::type;
    constexpr int U = sizeof(T) == 4 ? 16 : 8;     // words in flight (measured best per type)
    const int64_t g0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * V;
    if (g0 >= (int64_t)N * P) return;
    const size_t stride = (size_t)N * P / V;
    const W* f = reinterpret_cast<const W*>(vdf) + g0 / V;
    double ss[V], ssx[V], ssy[V];
#pragma unroll
    for (int e = 0; e < V; ++e) ss[e] = ssx[e] = ssy[e] = 0.0;
    auto acc = [&](const W& x, int d) {
        const T* t = reinterpret_cast<const T*>(&x);
        const double w = dw[d], cx = cxv[d], cy = cyv[d];
#pragma unroll
        for (int e = 0; e < V; ++e) {
            const double v = (double)t[e];
            ss[e] += v * w;
            ssx[e] += cx * v * w;
            ssy[e] += cy * v * w;
        }
    };
    int d = 0;
    for (; d + U <= D; d += U) {
        W x[U];
#pragma unroll
        for (int u = 0; u < U; ++u) x[u] = f[(size_t)(d + u) * stride];
#pragma unroll
        for (int u = 0; u < U; ++u) acc(x[u], d + u);
    }
    for (; d < D; ++d) acc(f[(size_t)d * stride], d);
#pragma unroll
    for (int e = 0; e < V; ++e) {
        const int64_t g = g0 + e;
        const int i = (int)(g / P), l = (int)(g - (int64_t)i * P);
        ts[(size_t)l * N + i] = ss[e] / cv;
        qxs[(size_t)l * N + i] = ssx[e];
        qys[(size_t)l * N + i] = ssy[e];
        ssraw[(size_t)l * N + i] = ss[e];
    }
}

__global__ void __launch_bounds__(256) integrals_kernel(const double* ssraw, const double* qxs,
                                                        const double* qys, const double* int_tri,
                                                        int N, int P, double cv, double* temp,
                                                        double* qx, double* qy)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    double t = 0.0, x = 0.0, y = 0.0;
    for (int l = 0; l < P; ++l) {
        const double w = int_tri[(size_t)l * N + i];
        t += ssraw[(size_t)l * N + i] * w / cv;
        x += qxs[(size_t)l * N + i] * w;
        y += qys[(size_t)l * N + i] * w;
    }
    temp[i] = t;
    qx[i] = x;
    qy[i] = y;
}

// ==========================================================================
// adiabatic wall emission (_kernels.py:363-397): one thread per (wall face, m)
// ==========================================================================
// sum over the 64 lanes of a wave, butterfly order (fixed, so the result
// does not depend on the launch)
__device__ inline double wave_sum64(double v)
{
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// Diffuse adiabatic wall emission (_kernels.py:364-397): one wave per wall
// face, lanes over directions (each lane its directions d = lane + 64 j in
// order), the lane partial sums combined by the butterfly; the P moments of
// a face share each lane's loads of the element's values.
template <typename T, int P>
__global__ void __launch_bounds__(64) wall_flux_kernel(const T* __restrict__ vdf, const double* cxv,
                                                       const double* cyv, const double* dw,
                                                       const double* fcm, const double* nx3,
                                                       const int32_t* wall_face, const int32_t* wall_tri,
                                                       const int32_t* wall_lfc, int D, int N,
                                                       double* flux_wall)
{
    const int k = blockIdx.x, lane = threadIdx.x;
    const int tri = wall_tri[k], lfc = wall_lfc[k];
    const double n1 = -nx3[2 * (tri * 3 + lfc)], n2 = -nx3[2 * (tri * 3 + lfc) + 1];
    const double* fc = fcm + ((size_t)tri * 3 + lfc) * P * P;
    double ss1[P], ss2 = 0.0;
#pragma unroll
    for (int m = 0; m < P; ++m) ss1[m] = 0.0;
    for (int d = lane; d < D; d += 64) {
        const double u = cxv[d] * n1 + cyv[d] * n2;
        if (u < 0.0) {
            const T* f = vdf + ((size_t)d * N + tri) * P;
            double fv[P];
#pragma unroll
            for (int l = 0; l < P; ++l) fv[l] = (double)f[l];
#pragma unroll
            for (int m = 0; m < P; ++m) {
                double acc = 0.0;
#pragma unroll
                for (int l = 0; l < P; ++l) acc += fv[l] * fc[m * P + l];
                ss1[m] += u * dw[d] * acc;
            }
        } else {
            ss2 -= u * dw[d];
        }
    }
    ss2 = wave_sum64(ss2);
#pragma unroll
    for (int m = 0; m < P; ++m) {
        const double a = wave_sum64(ss1[m]);
        if (lane == 0) flux_wall[(size_t)wall_face[k] * P + m] = -a / ss2;
    }
}

// ==========================================================================
// transport residual ||A f - b||^2 (_kernels.py:274-306), binary64, per
// direction.  A workgroup covers 256 cells and 64 directions, each wave 64
// consecutive cells in turn with one lane per direction, so the element
// tensors are wave-uniform (scalar) loads, read once per 64 directions, and
// only the distribution rows are per lane.  Per direction a lane sums its
// 64 cells in index order, then the four waves are added in wave order:
// partial[d][block][2] (numerator, denominator).
// ==========================================================================
template <int P>
__global__ void __launch_bounds__(256) residual_cells_kernel(
    const double* __restrict__ vdf, const double* __restrict__ src, const double* __restrict__ cxv,
    const double* __restrict__ cyv, const double* __restrict__ dw, const double* __restrict__ mass,
    const double* __restrict__ gx, const double* __restrict__ gy, const double* __restrict__ fcm,
    const CellGeo<double>* __restrict__ geo, const double* __restrict__ K, const double* __restrict__ bvec,
    const double* __restrict__ flux_wall, double inv_tau_c, int D, int N, double* __restrict__ partial)
{
    constexpr int DEG = Deg<P>::v;
    constexpr int Q = DEG + 1;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int d = blockIdx.y * 64 + lane;
    const int dd = d < D ? d : D - 1;
    const double cx = cxv[dd], cy = cyv[dd], wd = dw[dd];
    const double* vd = vdf + (size_t)dd * N * P;
    double num = 0.0, den = 0.0;
    const int c0 = __builtin_amdgcn_readfirstlane(blockIdx.x * 256 + wave * 64);
    const int c1 = min(c0 + 64, N);
    for (int i = c0; i < c1; ++i) {
        double b[P], f[P];
        const double* sv = src + (size_t)i * Lay<P>::SRC;
#pragma unroll
        for (int m = 0; m < P; ++m) {
            b[m] = sv[m] + cx * sv[P + m] + cy * sv[2 * P + m];
            f[m] = vd[(size_t)i * P + m];
        }
        const CellGeo<double> g = geo[i];
        double wout[3];
#pragma unroll
        for (int il = 0; il < 3; ++il) {
            const double sp = cx * g.n[2 * il] + cy * g.n[2 * il + 1];
            wout[il] = 0.5 * (sp + fabs(sp));
            const double w = 0.5 * (sp - fabs(sp));
            const int info = g.info[il];
            const int code = info & 3;
            if (code <= FACE_PERIODIC) {
                const int ilp = (info >> 2) & 3;
                const double* fe = vd + (size_t)g.other[il] * P;
                const double* Kf = K + ((size_t)i * 3 + il) * Lay<P>::KQ;
                if (w != 0.0) {
#pragma unroll
                    for (int r = 0; r < Q; ++r) {
                        double acc = b[face_node(DEG, il, r)];
#pragma unroll
                        for (int q = 0; q < Q; ++q) acc -= (w * Kf[r * Q + q]) * fe[face_node(DEG, ilp, q)];
                        b[face_node(DEG, il, r)] = acc;
                    }
                }
            } else if (code == FACE_ADIABATIC) {
                if (w != 0.0) {
#pragma unroll
                    for (int m = 0; m < P; ++m) b[m] += w * flux_wall[(size_t)g.other[il] * P + m];
                }
            } else if (w != 0.0) {
#pragma unroll
                for (int m = 0; m < P; ++m) b[m] -= w * bvec[(size_t)g.other[il] * pad4(P) + m];
            }
        }
        const double* fc = fcm + (size_t)i * 3 * P * P;
#pragma unroll
        for (int m = 0; m < P; ++m) {
            double r = -b[m];
#pragma unroll
            for (int l = 0; l < P; ++l) {
                double a = (inv_tau_c * mass[((size_t)m * P + l) * N + i] - cx * gx[((size_t)i * P + m) * P + l] -
                            cy * gy[((size_t)i * P + m) * P + l]);
                a = a + wout[0] * fc[(0 * P + m) * P + l];
                a = a + wout[1] * fc[(1 * P + m) * P + l];
                a = a + wout[2] * fc[(2 * P + m) * P + l];
                r += a * f[l];
            }
            num += r * r * wd;
            den += b[m] * b[m] * wd;
        }
    }
    __shared__ double sn[4][64], sd[4][64];
    sn[wave][lane] = num;
    sd[wave][lane] = den;
    __syncthreads();
    if (wave == 0 && d < D) {
        double a = sn[0][lane], c = sd[0][lane];
#pragma unroll
        for (int k = 1; k < 4; ++k) {
            a += sn[k][lane];
            c += sd[k][lane];
        }
        partial[((size_t)d * gridDim.x + blockIdx.x) * 2 + 0] = a;
        partial[((size_t)d * gridDim.x + blockIdx.x) * 2 + 1] = c;
    }
}

// fixed-order block reduction: thread t sums elements t, t + 256, ... in
// index order, then a pairwise tree over the 256 partials.  The order never
// depends on timing, so every run gives the same bits.
template <int NV>
__device__ __forceinline__ void block_sum256(double (&v)[NV], double (*sh)[256])
{
    const int t = threadIdx.x;
#pragma unroll
    for (int j = 0; j < NV; ++j) sh[j][t] = v[j];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
#pragma unroll
            for (int j = 0; j < NV; ++j) sh[j][t] = sh[j][t] + sh[j][t + w];
        }
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < NV; ++j) v[j] = sh[j][0];
}

// out[g] = sum over k of in[(r * ncol + k) * stride_k + c], g = r * comps + c:
// one workgroup per output
__global__ void __launch_bounds__(256) rowsum_kernel(const double* in, int nrow, int ncol, int stride_k, int comps,
                                                     double* out)
{
    __shared__ double sh[1][256];
    const int g = blockIdx.x;
    const int r = g / comps, c = g - (g / comps) * comps;
    double v[1] = {0.0};
    for (int k = threadIdx.x; k < ncol; k += 256) v[0] += in[((size_t)r * ncol + k) * stride_k + c];
    block_sum256<1>(v, sh);
    if (threadIdx.x == 0) out[g] = v[0];
}

// residual and mass of an iterate (_kernels.py:340-361) in two passes: per
// workgroup partial sums over a grid-stride partition, then one workgroup
// sums the partials in block order.  The order is fixed (a run repeats
// bitwise), but it is not the reference's sequential index order.
constexpr int kIterBlocks = 64;

__global__ void __launch_bounds__(256) iterate_partial_kernel(const double* temp, const double* t_old, int n,
                                                              double* partial)
{
    __shared__ double sh[3][256];
    double v[3] = {0.0, 0.0, 0.0};
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const double dlt = temp[i] - t_old[i];
        v[0] += dlt * dlt;
        v[1] += temp[i] * temp[i];
        v[2] += temp[i];
    }
    block_sum256<3>(v, sh);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) partial[k * gridDim.x + blockIdx.x] = v[k];
    }
}

__global__ void __launch_bounds__(256) iterate_residual_kernel(const double* partial, int nbk, double* out)
{
    __shared__ double sh[3][256];
    double v[3] = {0.0, 0.0, 0.0};
    if ((int)threadIdx.x < nbk) {
#pragma unroll
        for (int k = 0; k < 3; ++k) v[k] = partial[k * nbk + threadIdx.x];
    }
    block_sum256<3>(v, sh);
    if (threadIdx.x == 0) {
        out[0] = sqrt(v[0] / v[1]);
        out[1] = v[2];
    }
}

// ==========================================================================
// Krylov vector kernels: fixed partition, fixed reduction order
// ==========================================================================
constexpr int kRows = 8;
constexpr int kDotBlocks = 1024;
constexpr int kMaxBasis = 512;   // rows of one multidot / update call

__global__ void __launch_bounds__(256) multidot_kernel(const double* __restrict__ Qm, int64_t ld, int m,
                                                       const double* __restrict__ w, int64_t n,
                                                       double* partial)
{
    const int r0 = blockIdx.y * kRows;
    double acc[kRows];
#pragma unroll
    for (int r = 0; r < kRows; ++r) acc[r] = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double wi = w[i];
#pragma unroll
        for (int r = 0; r < kRows; ++r)
            if (r0 + r < m) acc[r] += Qm[(size_t)(r0 + r) * ld + i] * wi;
    }
    __shared__ double sh[kRows][256];
#pragma unroll
    for (int r = 0; r < kRows; ++r) sh[r][threadIdx.x] = acc[r];
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
#pragma unroll
            for (int r = 0; r < kRows; ++r) sh[r][threadIdx.x] += sh[r][threadIdx.x + h];
        __syncthreads();
    }
    if ((int)threadIdx.x < kRows && r0 + (int)threadIdx.x < m)
        partial[(size_t)(r0 + threadIdx.x) * gridDim.x + blockIdx.x] = sh[threadIdx.x][0];
}

__global__ void __launch_bounds__(256) update_kernel(const double* __restrict__ Qm, int64_t ld, int m,
                                                     const double* __restrict__ c, double* w, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int k = 0; k < m; ++k) s += c[k] * Qm[(size_t)k * ld + i];
    w[i] -= s;
}

// w -= Q^T c with update_kernel's arithmetic per element, then, over the
// same pass, the dot products of the updated w with the m rows of Q
// (NORM: with itself) with multidot_kernel's partition, per-thread order and
// tree, so the partials are those of an update followed by a multidot bit
// for bit, for one read of Q instead of two.  Launched on multidot's
// gridDim.x blocks; m <= MR.
template <int MR, bool NORM>
__global__ void __launch_bounds__(256) update_dot_kernel(const double* __restrict__ Qm, int64_t ld, int m,
                                                         const double* __restrict__ c, double* __restrict__ w,
                                                         int64_t n, double* __restrict__ partial)
{
    constexpr int NR = NORM ? 1 : MR;
    double acc[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        // the element's column of Q, all loads issued before the sums
        double q[MR];
#pragma unroll
        for (int k = 0; k < MR; ++k) q[k] = k < m ? Qm[(size_t)k * ld + i] : 0.0;
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < MR; ++k)
            if (k < m) s += c[k] * q[k];
        const double wn = w[i] - s;
        w[i] = wn;
        if constexpr (NORM) {
            acc[0] += wn * wn;
        } else {
#pragma unroll
            for (int r = 0; r < MR; ++r)
                if (r < m) acc[r] += q[r] * wn;
        }
    }
    __shared__ double sh[kRows][256];
    const int rows = NORM ? 1 : m;
#pragma unroll
    for (int g = 0; g < (NR + kRows - 1) / kRows; ++g) {
        if (g * kRows >= rows) break;
#pragma unroll
        for (int r = 0; r < kRows; ++r) sh[r][threadIdx.x] = g * kRows + r < NR ? acc[g * kRows + r] : 0.0;
        __syncthreads();
        for (int h = blockDim.x / 2; h > 0; h >>= 1) {
            if ((int)threadIdx.x < h)
#pragma unroll
                for (int r = 0; r < kRows; ++r) sh[r][threadIdx.x] += sh[r][threadIdx.x + h];
            __syncthreads();
        }
        if ((int)threadIdx.x < kRows && g * kRows + (int)threadIdx.x < rows)
            partial[(size_t)(g * kRows + threadIdx.x) * gridDim.x + blockIdx.x] = sh[threadIdx.x][0];
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) axpby_kernel(double a, const double* x, double b, double* y, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    y[i] = a * x[i] + b * y[i];
}

// y = (1 / |w|) w with |w|^2 read from the device: axpby_kernel's
// expression with a = 1 / sqrt(max(|w|^2, 0)) and b = 0
__global__ void __launch_bounds__(256) normalize_kernel(const double* __restrict__ n2, const double* x, double* y,
                                                        int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v = n2[0];
    const double a = 1.0 / sqrt(v > 0.0 ? v : 0.0), b = 0.0;
    y[i] = a * x[i] + b * y[i];
}

// out = (v - t) + g, the order of out = v; out += -t; out += g
__global__ void __launch_bounds__(256) fixed_point_defect_kernel(const double* v, const double* t, const double* g,
                                                                 double* out, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double o = -1.0 * t[i] + 1.0 * v[i];
    if (g) o = 1.0 * g[i] + 1.0 * o;
    out[i] = o;
}

// max of two magnitudes where a NaN wins, as numpy's max does
__device__ inline double nan_max(double a, double b) { return (a != a || a > b) ? a : b; }

// The relaxation of the GSIS fixed-point repair
// (acceleration/__init__.py:144-147): per workgroup the NaN-propagating max
// of |o - d| over a grid-stride partition, and, when omega != 0,
// d += omega (o - d) in numpy's rounding (difference, product, sum, each
// rounded; separate statements, so nothing is contracted into an fma).
__global__ void __launch_bounds__(256) relax_kernel(double omega, const double* __restrict__ o, double* d,
                                                    int64_t n, double* partial)
{
    double m = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const double di = d[i];
        const double t = o[i] - di;
        m = nan_max(fabs(t), m);
        if (omega != 0.0) {
            const double u = omega * t;
            d[i] = di + u;
        }
    }
    __shared__ double sh[256];
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) sh[threadIdx.x] = nan_max(sh[threadIdx.x], sh[threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(256) max_final_kernel(const double* partial, int nbk, double* out)
{
    __shared__ double sh[256];
    double m = 0.0;
    for (int k = threadIdx.x; k < nbk; k += 256) m = nan_max(partial[k], m);
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) sh[threadIdx.x] = nan_max(sh[threadIdx.x], sh[threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

// periodic partner DOFs as Krylov state: gather vdf[d][ptri[k]][:] into a
// binary64 block [D][K][P], and scatter such a block into pbuf[d][pface[k]][:]
template <typename T>
__global__ void __launch_bounds__(256) periodic_gather_kernel(const T* __restrict__ vdf, const int32_t* ptri,
                                                              int K, int D, int N, int P, double* out)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (int64_t)D * K * P) return;
    const int l = (int)(g % P);
    const int k = (int)((g / P) % K);
    const int d = (int)(g / ((int64_t)P * K));
    out[g] = (double)vdf[((size_t)d * N + ptri[k]) * P + l];
}

template <typename T>
__global__ void __launch_bounds__(256) periodic_scatter_kernel(const double* __restrict__ in, const int32_t* pface,
                                                               int K, int D, int F, int P, T* pbuf)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (int64_t)D * K * P) return;
    const int l = (int)(g % P);
    const int k = (int)((g / P) % K);
    const int d = (int)(g / ((int64_t)P * K));
    pbuf[((size_t)d * F + pface[k]) * P + l] = (T)in[g];
}

// ==========================================================================
// GSIS with the HDG synthetic system (acceleration/*.py), per iteration:
//   moments7 -> hot work -> AA_SRC -> trace RHS (+ wall traces) -> [trace
//   solve] -> local recovery -> kinetic correction
// ==========================================================================

// seven angular moments per (cell, dof), binary64, directions in index
// order: 1, cx, cy (as moments_kernel) and the four third-order weights of
// the higher-order-term source, cx(5cx^2-3), cy(5cx^2-1), cx(5cy^2-1),
// cy(5cy^2-3), each times dw (hot_source.py:29-58 summed over l first)
template <typename T, int V>
__global__ void __launch_bounds__(256) moments7_kernel(const T* __restrict__ vdf, const double* __restrict__ cxv,
                                                       const double* __restrict__ cyv,
                                                       const double* __restrict__ dw,
                                                       const double* __restrict__ hw, int D, int N, int P,
                                                       double* __restrict__ out)
{
    // V consecutive (cell, dof) values per thread, one 16-byte word per
    // direction (V = 1: one value), U words in flight; each value's seven
    // sums run over the directions in index order
    using W = typename Vec16<T>::type;
    constexpr int U = sizeof(T) == 4 ? 8 : 4;
    const int64_t g0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * V;
    if (g0 >= (int64_t)N * P) return;
    double a[7][V];
#pragma unroll
    for (int k = 0; k < 7; ++k)
#pragma unroll
        for (int e = 0; e < V; ++e) a[k][e] = 0.0;
    auto acc = [&](const T* t, int d) {
        const double w = dw[d], cx = cxv[d], cy = cyv[d];
        const double h0 = hw[d], h1 = hw[D + d], h2 = hw[2 * D + d], h3 = hw[3 * D + d];
#pragma unroll
        for (int e = 0; e < V; ++e) {
            const double f = (double)t[e];
            a[0][e] += f * w;
            a[1][e] += cx * f * w;
            a[2][e] += cy * f * w;
            a[3][e] += h0 * f;
            a[4][e] += h1 * f;
            a[5][e] += h2 * f;
            a[6][e] += h3 * f;
        }
    };
    const size_t stride = (size_t)N * P;
    if constexpr (V * sizeof(T) == 16) {
        const W* f = reinterpret_cast<const W*>(vdf) + g0 / V;
        const size_t ws = stride / V;
        int d = 0;
        for (; d + U <= D; d += U) {
            W x[U];
#pragma unroll
            for (int u = 0; u < U; ++u) x[u] = f[(size_t)(d + u) * ws];
#pragma unroll
            for (int u = 0; u < U; ++u) acc(reinterpret_cast<const T*>(&x[u]), d + u);
        }
        for (; d < D; ++d) {
            const W x = f[(size_t)d * ws];
            acc(reinterpret_cast<const T*>(&x), d);
        }
    } else {
        for (int d = 0; d < D; ++d) {
            const T x = vdf[(size_t)d * stride + g0];
            acc(&x, d);
        }
    }
    const size_t PN = (size_t)P * N;
#pragma unroll
    for (int e = 0; e < V; ++e) {
        const int64_t g = g0 + e;
        const int i = (int)(g / P), l = (int)(g - (int64_t)i * P);
        const size_t o = (size_t)l * N + i;
#pragma unroll
        for (int k = 0; k < 7; ++k) out[k * PN + o] = a[k][e];
    }
}

// GSIS element operators.  With w the stress work of an element (4P
// values, the HoT source's stress rows) the accelerated step needs, per
// element, the rows T, qx, qy of AA_SRC = inv(AA_SOL)[:, 3P:7P] w and the
// three per-side trace sums BA_SOL[l] AA_SRC (hot_source.py:61-75,
// global_solve.py:128-150); both are linear in w, so one operator
// op[i][q][k] (k < 3P: inverse rows, k >= 3P: (l, r) = divmod(k - 3P, 3Q),
// sum_p BA_SOL[i][l][r][p] inv[i][p][c0 + q], p in order) maps w to all
// K = 3P + 9Q outputs, and the local recovery needs only AA_TRACE's rows
// of T, qx, qy (local_solve.py:16-26; correction.py reads those three).
// Built from the columns c0 .. c0 + W of the inverse: the stress columns
// (c0 = 3P, W = 4P) for the step, the T, qx, qy columns (c0 = 0, W = 3P)
// for the moment preconditioner (krylov.py:388-402).
__global__ void gsis_op_kernel(const double* __restrict__ inv, const double* __restrict__ ba, int64_t N, int P,
                               int Q, int c0, int W, double* __restrict__ out)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int R = 7 * P, K = 3 * P + 9 * Q;
    const int64_t per = (int64_t)W * K;
    if (g >= N * per) return;
    const int64_t i = g / per;
    const int e = (int)(g - i * per);
    const int q = e / K, k = e - (e / K) * K;
    const double* iv = inv + (size_t)i * R * R;
    if (k < 3 * P) {
        out[g] = iv[(size_t)k * R + c0 + q];
        return;
    }
    const int kk = k - 3 * P, l = kk / (3 * Q), r = kk - (kk / (3 * Q)) * (3 * Q);
    const double* b = ba + (((size_t)i * 3 + l) * 3 * Q + r) * R;
    double s = 0.0;
    for (int pp = 0; pp < R; ++pp) s += b[pp] * iv[(size_t)pp * R + c0 + q];
    out[g] = s;
}

// rows r0 .. r0 + nr of the stress columns of the inverse, out[i][q][k] =
// inv[i][r0 + k][3P + q] (the AA_SRC rows 3P .. 7P of the full recovery)
__global__ void inv_rows_kernel(const double* __restrict__ inv, int64_t N, int P, int r0, int nr,
                                double* __restrict__ out)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int R = 7 * P, W = 4 * P;
    const int64_t per = (int64_t)W * nr;
    if (g >= N * per) return;
    const int64_t i = g / per;
    const int e = (int)(g - i * per);
    const int q = e / nr, k = e - (e / nr) * nr;
    out[g] = inv[((size_t)i * R + r0 + k) * R + 3 * P + q];
}

// rows p0 .. p0 + np of AA_TRACE [N][3][7P][3Q] as out[i][l][r][p - p0]
__global__ void aat_rows_kernel(const double* __restrict__ aat, int64_t N, int P, int Q, int p0, int np,
                                double* __restrict__ out)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t per = (int64_t)9 * Q * np;
    if (g >= N * per) return;
    const int64_t i = g / per;
    const int e = (int)(g - i * per);
    const int lr = e / np, p = e - (e / np) * np;
    const int l = lr / (3 * Q), r = lr - (lr / (3 * Q)) * (3 * Q);
    out[g] = aat[(((size_t)i * 3 + l) * 7 * P + p0 + p) * 3 * Q + r];
}

// the fused sweep's element scalar area2 / tau_c from the cell operators
// [N][8] (area2, the gradient coefficients (y3-y1)/s, (y1-y2)/s, (x1-x3)/s,
// (x2-x1)/s, ...) into the cell records' padding at column rf, both
// precisions; and the check that area2 times the gradient coefficients are
// minus the length-scaled normals the sweep reads at column rn (faces 2 and
// 0), counting elements off by more than 1e-10 of the largest in *bad
__global__ void __launch_bounds__(256) fused_scalars_kernel(const double* __restrict__ cell, double tau_c, int N,
                                                            int rf, int rn, double* crec64, int rs64, float* crec32,
                                                            int rs32, int32_t* bad)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double* c = cell + (size_t)i * 8;
    const double am = c[0] / tau_c;
    crec64[(size_t)i * rs64 + rf] = am;
    crec32[(size_t)i * rs32 + rf] = (float)am;
    const double* nl = crec64 + (size_t)i * rs64 + rn;
    const double g[4] = {c[0] * c[1], c[0] * c[3], c[0] * c[2], c[0] * c[4]};
    const double n[4] = {nl[4], nl[5], nl[0], nl[1]};
    double big = 0.0, err = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        big = fmax(big, fabs(n[k]));
        err = fmax(err, fabs(g[k] + n[k]));
    }
    if (!(err <= 1e-10 * big)) atomicAdd(bad, 1);
}

template <typename T>
__global__ void narrow_kernel(const double* __restrict__ in, T* __restrict__ out, int64_t n)
{
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < n; g += (int64_t)gridDim.x * blockDim.x)
        out[g] = (T)in[g];
}

// the stress work of element i (hot_source.py:29-58): w[q], q < 4P, for
// dof m, the l sums in order
template <int P>
__device__ inline void stress_work(const double* __restrict__ m7, const double* __restrict__ gx,
                                   const double* __restrict__ gy, int64_t N, int64_t i, int m, double tc5,
                                   double* w)
{
    const size_t PN = (size_t)P * N;
    double pixx = 0, pixy = 0, piyy = 0;
    for (int l = 0; l < P; ++l) {
        const double ax = gx[((size_t)i * P + l) * P + m], ay = gy[((size_t)i * P + l) * P + m];
        const double w1 = m7[3 * PN + (size_t)l * N + i], w2 = m7[4 * PN + (size_t)l * N + i];
        const double w3 = m7[5 * PN + (size_t)l * N + i], w4 = m7[6 * PN + (size_t)l * N + i];
        pixx += ax * w1 + ay * w2;
        pixy += ax * w2 + ay * w3;
        piyy += ax * w3 + ay * w4;
    }
    w[m] = (pixx + 0.5 * piyy) * tc5;
    w[P + m] = 0.5 * pixy * tc5;
    w[2 * P + m] = 0.5 * pixy * tc5;
    w[3 * P + m] = (0.5 * pixx + piyy) * tc5;
}

// the element outputs of one workgroup's EB elements from their inputs in
// LDS (xin[e][q], q < W): src[k][i] (k < 3P) and side[i][k - 3P] (k >= 3P),
// s = sum_q op[i][q][k] xin[q], q in order; threads over (element, k), so
// that consecutive threads read consecutive operator words
template <int EB, typename TO>
__device__ inline void element_apply(const TO* __restrict__ op, const double (*xin)[64], int64_t N, int64_t i0,
                                     int P, int Q, int W, double* __restrict__ src, double* __restrict__ side)
{
    const int K = 3 * P + 9 * Q;
    for (int t = threadIdx.x; t < EB * K; t += blockDim.x) {
        const int e = t / K, k = t - (t / K) * K;
        const int64_t i = i0 + e;
        if (i >= N) continue;
        const TO* a = op + (size_t)i * W * K + k;
        double s = 0.0;
        for (int q = 0; q < W; ++q) s += (double)a[(size_t)q * K] * xin[e][q];
        if (k < 3 * P)
            src[(size_t)k * N + i] = s;
        else
            side[(size_t)i * 9 * Q + k - 3 * P] = s;
    }
}

// the accelerated step's element pass: the stress work from the third-order
// moments, then AA_SRC's T, qx, qy rows and the side sums through op
template <int P, int Q, typename TO>
__global__ void __launch_bounds__(256) hot_side_kernel(const double* __restrict__ m7, const double* __restrict__ gx,
                                                       const double* __restrict__ gy, const TO* __restrict__ op,
                                                       int64_t N, double tc5, double* __restrict__ src,
                                                       double* __restrict__ side)
{
    constexpr int EB = 8;
    static_assert(4 * P <= 64, "stress work exceeds the LDS row");
    __shared__ double w[EB][64];
    const int64_t i0 = (int64_t)blockIdx.x * EB;
    for (int t = threadIdx.x; t < EB * P; t += blockDim.x) {
        const int e = t / P, m = t - (t / P) * P;
        if (i0 + e < N) stress_work<P>(m7, gx, gy, N, i0 + e, m, tc5, w[e]);
    }
    __syncthreads();
    element_apply<EB, TO>(op, w, N, i0, P, Q, 4 * P, src, side);
}

// the moment preconditioner's element pass: sign times the T, qx, qy
// blocks of r (three [P][N] blocks) through the preconditioner's operator
template <typename TO>
__global__ void __launch_bounds__(256) precond_side_kernel(const double* __restrict__ r, double sign,
                                                           const TO* __restrict__ op, int64_t N, int P, int Q,
                                                           double* __restrict__ src, double* __restrict__ side)
{
    constexpr int EB = 8;
    __shared__ double w[EB][64];
    const int64_t i0 = (int64_t)blockIdx.x * EB;
    for (int t = threadIdx.x; t < EB * 3 * P; t += blockDim.x) {
        const int e = t / (3 * P), q = t - (t / (3 * P)) * (3 * P);
        if (i0 + e < N) w[e][q] = sign * r[(size_t)q * N + i0 + e];
    }
    __syncthreads();
    element_apply<EB, TO>(op, w, N, i0, P, Q, 3 * P, src, side);
}


// wall trace moments (global_solve.py:31-53): one wave per wall face,
// lanes over directions (d = lane + 64 j in order), the Q face nodes' three
// moments combined over lanes by the butterfly
template <typename T, int P, int Q>
__global__ void __launch_bounds__(64) wall_trace_kernel(const T* __restrict__ vdf, const double* cxv, const double* cyv,
                                                        const double* dw, const double* __restrict__ wtfc,
                                                        const int32_t* wtri, int D, int N, double* out)
{
    const int k = blockIdx.x, lane = threadIdx.x;
    const int tri = wtri[k];
    const double* c = wtfc + (size_t)k * Q * P;
    double tb[Q], qxb[Q], qyb[Q];
#pragma unroll
    for (int m = 0; m < Q; ++m) tb[m] = qxb[m] = qyb[m] = 0.0;
    for (int d = lane; d < D; d += 64) {
        const T* f = vdf + ((size_t)d * N + tri) * P;
        double fv[P];
#pragma unroll
        for (int l = 0; l < P; ++l) fv[l] = (double)f[l];
#pragma unroll
        for (int m = 0; m < Q; ++m) {
            double t = 0.0;
#pragma unroll
            for (int l = 0; l < P; ++l) t += fv[l] * c[m * P + l];
            tb[m] += t * dw[d];
            qxb[m] += cxv[d] * t * dw[d];
            qyb[m] += cyv[d] * t * dw[d];
        }
    }
#pragma unroll
    for (int m = 0; m < Q; ++m) {
        const double a = wave_sum64(tb[m]), b = wave_sum64(qxb[m]), e = wave_sum64(qyb[m]);
        if (lane == 0) {
            out[(size_t)k * 3 * Q + m] = a;
            out[(size_t)k * 3 * Q + Q + m] = b;
            out[(size_t)k * 3 * Q + 2 * Q + m] = e;
        }
    }
}

// trace right-hand side (global_solve.py:128-186) from the per-side sums
// side[t][l][r] of the element pass (hot_side_kernel, precond_side_kernel):
// then per (face, row): ffa[f][r] = side(t1, l1)[r] + side(t2, l2)[r'], r'
// mirrored within each field block across a reversed periodic pair
__global__ void __launch_bounds__(256) trace_rhs_kernel(const double* __restrict__ side, const int32_t* t1,
                                                        const int32_t* l1, const int32_t* t2, const int32_t* l2,
                                                        const uint8_t* flip, int F, int Q, double* ffa)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    const int R = 3 * Q;
    if (g >= F * R) return;
    const int f = g / R, r = g - (g / R) * R;
    double s = side[((size_t)t1[f] * 3 + l1[f]) * R + r];
    if (t2[f] >= 0) {
        int rr = r;
        if (flip[f]) {   // mirror_permutation: k -> Q-1-k within each field block
            const int blk = r / Q, k = r - blk * Q;
            rr = blk * Q + (Q - 1 - k);
        }
        s += side[((size_t)t2[f] * 3 + l2[f]) * R + rr];
    }
    ffa[(size_t)f * R + r] = s;
}

// walls impose the trace from the kinetic solution (global_solve.py:151-186)
__global__ void __launch_bounds__(256) trace_wall_kernel(const double* __restrict__ mom, const int32_t* wface,
                                                         const double* wnx, const double* wny, int nw, int n_th,
                                                         int Q, double cv, int variant_b, double* ffa)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nw * Q) return;
    const int k = g / Q, m = g - (g / Q) * Q;
    const int R = 3 * Q;
    double* o = ffa + (size_t)wface[k] * R;
    const double* w = mom + (size_t)k * R;
    o[m] = w[m] / cv;
    if (k < n_th) {
        if (variant_b) {
            o[Q + m] = w[Q + m];
            o[2 * Q + m] = w[2 * Q + m];
        }
    } else {
        const double qxb = w[Q + m], qyb = w[2 * Q + m], nx = wnx[k], ny = wny[k];
        o[Q + m] = qxb * (ny * ny) - qyb * (nx * ny);
        o[2 * Q + m] = -qxb * (nx * ny) + qyb * (nx * nx);
    }
}

// y[0:n] = v (a fill that a captured graph can hold, unlike a memset)
__global__ void __launch_bounds__(256) fill_kernel(double* y, int64_t n, double v)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n) y[g] = v;
}

// UQ's T, qx, qy rows: uq[p][i] = src[p][i] + sum_il AA_TRACE[i][il][p]
// U_TRACE[face(i, il)] (local_solve.py:16-26), il then r in order, one
// thread per (element, p) from aat [N][3][3Q][3P]
template <typename TO>
__global__ void __launch_bounds__(256) local3_kernel(const TO* __restrict__ aat, const double* __restrict__ src,
                                                     const double* __restrict__ u, const int32_t* __restrict__ tri_faces,
                                                     int64_t N, int P, int Q, double* __restrict__ uq)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int R = 3 * P, C = 3 * Q;
    if (g >= N * R) return;
    const int64_t i = g / R;
    const int p = (int)(g - i * R);
    double s = 0.0;
    for (int il = 0; il < 3; ++il) {
        const TO* a = aat + ((size_t)i * 3 + il) * C * R + p;
        const double* x = u + (size_t)tri_faces[i * 3 + il] * C;
        for (int r = 0; r < C; ++r) s += (double)a[(size_t)r * R] * x[r];
    }
    uq[(size_t)p * N + i] = src[(size_t)p * N + i] + s;
}

// the full UQ [7P][N] of a step (its stress work recomputed from m7): rows
// below 3P as local3_kernel, rows from 3P through op_hi [N][4P][4P] and
// aat_hi [N][3][3Q][4P]; the field output reads the stress rows
template <int P, int Q>
__global__ void __launch_bounds__(256) uq_full_kernel(const double* __restrict__ m7, const double* __restrict__ gx,
                                                      const double* __restrict__ gy, const double* __restrict__ op,
                                                      const double* __restrict__ op_hi,
                                                      const double* __restrict__ aat, const double* __restrict__ aat_hi,
                                                      const double* __restrict__ u,
                                                      const int32_t* __restrict__ tri_faces, int64_t N, double tc5,
                                                      double* __restrict__ uq)
{
    constexpr int K = 3 * P + 9 * Q, W = 4 * P, C = 3 * Q;
    __shared__ double w[4][64];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t i = (int64_t)blockIdx.x * 4 + wv;
    if (i < N && lane < P) stress_work<P>(m7, gx, gy, N, i, lane, tc5, w[wv]);
    __syncthreads();
    if (i >= N) return;
    for (int k = lane; k < 7 * P; k += 64) {
        const bool lo = k < 3 * P;
        const int kk = lo ? k : k - 3 * P, stride = lo ? K : W, np = lo ? 3 * P : W;
        const double* a = (lo ? op + (size_t)i * W * K : op_hi + (size_t)i * W * W) + kk;
        double s = 0.0;
        for (int q = 0; q < W; ++q) s += a[(size_t)q * stride] * w[wv][q];
        double t = 0.0;
        for (int il = 0; il < 3; ++il) {
            const double* b = (lo ? aat : aat_hi) + ((size_t)i * 3 + il) * C * np + kk;
            const double* x = u + (size_t)tri_faces[i * 3 + il] * C;
            for (int r = 0; r < C; ++r) t += b[(size_t)r * np] * x[r];
        }
        uq[(size_t)k * N + i] = s + t;
    }
}


// damped kinetic correction (correction.py:29-75) per (cell, dof)
template <typename T>
__global__ void __launch_bounds__(256) correct_kernel(T* vdf, const double* __restrict__ m7, const double* __restrict__ uq,
                                                      const double* __restrict__ defect, const double* cxv,
                                                      const double* cyv, const double* hmin, int D, int N, int P,
                                                      double cv, double pi, double tau_r, double tau_thr,
                                                      double drift, int use_hmin, double* ts, double* qxs,
                                                      double* qys, double* mkin)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= N * P) return;
    const int i = g / P, m = g - (g / P) * P;
    const double tau_loc = use_hmin ? tau_r / hmin[i] : tau_r;
    const double beta = fmin(tau_loc, tau_thr) / tau_loc;
    const size_t PN = (size_t)P * N, o = (size_t)m * N + i;
    const double t_vdf = m7[o] / cv, qx_vdf = m7[PN + o], qy_vdf = m7[2 * PN + o];
    if (mkin) {
        mkin[o] = t_vdf;
        mkin[PN + o] = qx_vdf;
        mkin[2 * PN + o] = qy_vdf;
    }
    const double d0 = defect ? defect[o] : 0.0, d1 = defect ? defect[PN + o] : 0.0,
                 d2 = defect ? defect[2 * PN + o] : 0.0;
    const double l_t = (uq[o] - d0 - t_vdf) * beta;
    const double l_qx = (uq[PN + o] - d1 - qx_vdf) * beta;
    const double l_qy = (uq[2 * PN + o] - d2 - qy_vdf) * beta;
    const double eq = l_t * cv / 4.0 / pi;
    const size_t stride = (size_t)N * P;
    for (int d = 0; d < D; ++d) {
        T* f = vdf + (size_t)d * stride + g;
        *f = (T)((double)*f + eq + (cxv[d] * l_qx + cyv[d] * l_qy) * drift);
    }
    ts[o] = l_t + t_vdf;
    qxs[o] = l_qx + qx_vdf;
    qys[o] = l_qy + qy_vdf;
}

// The moment half of the kinetic correction (correction.py:29-75): the
// corrected dofs, the kinetic moments mkin and, per (cell, dof), the
// correction coefficients coef[3][P][N] = (eq, l_qx, l_qy) that
// apply_correction_kernel adds to the distribution (the same expressions as
// correct_kernel, so both give the same bits)
__global__ void __launch_bounds__(256) correct_moments_kernel(const double* __restrict__ m7, const double* __restrict__ uq,
                                                              const double* __restrict__ defect,
                                                              const double* hmin, int N, int P, double cv, double pi,
                                                              double tau_r, double tau_thr, int use_hmin,
                                                              double* ts, double* qxs, double* qys, double* mkin,
                                                              double* coef)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= N * P) return;
    const int i = g / P, m = g - (g / P) * P;
    const double tau_loc = use_hmin ? tau_r / hmin[i] : tau_r;
    const double beta = fmin(tau_loc, tau_thr) / tau_loc;
    const size_t PN = (size_t)P * N, o = (size_t)m * N + i;
    const double t_vdf = m7[o] / cv, qx_vdf = m7[PN + o], qy_vdf = m7[2 * PN + o];
    if (mkin) {
        mkin[o] = t_vdf;
        mkin[PN + o] = qx_vdf;
        mkin[2 * PN + o] = qy_vdf;
    }
    const double d0 = defect ? defect[o] : 0.0, d1 = defect ? defect[PN + o] : 0.0,
                 d2 = defect ? defect[2 * PN + o] : 0.0;
    const double l_t = (uq[o] - d0 - t_vdf) * beta;
    const double l_qx = (uq[PN + o] - d1 - qx_vdf) * beta;
    const double l_qy = (uq[2 * PN + o] - d2 - qy_vdf) * beta;
    coef[o] = l_t * cv / 4.0 / pi;
    coef[PN + o] = l_qx;
    coef[2 * PN + o] = l_qy;
    ts[o] = l_t + t_vdf;
    qxs[o] = l_qx + qx_vdf;
    qys[o] = l_qy + qy_vdf;
}

// The distribution half: f += eq + (cx l_qx + cy l_qy) drift for every
// direction, on every cell (cells == NULL; V consecutive values per thread,
// one 16-byte word) or on the listed cells (V = 1, one (cell, dof) per thread)
template <typename T, int V>
__global__ void __launch_bounds__(256) apply_correction_kernel(T* vdf, const double* __restrict__ coef,
                                                               const double* cxv, const double* cyv,
                                                               const int32_t* __restrict__ cells, int ncells,
                                                               int D, int N, int P, double drift)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = cells ? (int64_t)ncells * P : (int64_t)N * P / V;
    if (t >= total) return;
    const size_t PN = (size_t)P * N, stride = PN;
    int64_t g0;
    if (cells) {
        const int k = (int)(t / P), m = (int)(t - (int64_t)k * P);
        g0 = (int64_t)cells[k] * P + m;
    } else {
        g0 = t * V;
    }
    double eq[V], lx[V], ly[V];
#pragma unroll
    for (int e = 0; e < V; ++e) {
        const int64_t g = g0 + e;
        const int i = (int)(g / P), m = (int)(g - (int64_t)i * P);
        const size_t o = (size_t)m * N + i;
        eq[e] = coef[o];
        lx[e] = coef[PN + o];
        ly[e] = coef[2 * PN + o];
    }
    for (int d = 0; d < D; ++d) {
        T* f = vdf + (size_t)d * stride + g0;
        T x[V];
#pragma unroll
        for (int e = 0; e < V; ++e) x[e] = f[e];
#pragma unroll
        for (int e = 0; e < V; ++e) x[e] = (T)((double)x[e] + eq[e] + (cxv[d] * lx[e] + cyv[d] * ly[e]) * drift);
#pragma unroll
        for (int e = 0; e < V; ++e) f[e] = x[e];
    }
}

// The same update on the listed cells, one thread per (direction, cell, dof):
// the lists are short (the elements behind adiabatic walls and periodic
// faces), so the directions are spread over threads rather than looped
template <typename T>
__global__ void __launch_bounds__(256) apply_correction_cells_kernel(T* vdf, const double* __restrict__ coef,
                                                                     const double* cxv, const double* cyv,
                                                                     const int32_t* __restrict__ cells, int ncells,
                                                                     int D, int N, int P, double drift)
{
    const int64_t per = (int64_t)ncells * P;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= per * D) return;
    const int d = (int)(t / per);
    const int64_t km = t - (int64_t)d * per;
    const int k = (int)(km / P), m = (int)(km - (int64_t)k * P);
    const int i = cells[k];
    const size_t PN = (size_t)P * N, o = (size_t)m * N + i;
    T* f = vdf + (size_t)d * PN + (size_t)i * P + m;
    *f = (T)((double)*f + coef[o] + (cxv[d] * coef[PN + o] + cyv[d] * coef[2 * PN + o]) * drift);
}

// element integrals of corrected dofs: temp[i] = sum_m ts[m][i] int_tri[m][i]
__global__ void __launch_bounds__(256) dof_integrals_kernel(const double* ts, const double* qxs, const double* qys,
                                                            const double* int_tri, int N, int P, double* temp,
                                                            double* qx, double* qy)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    double t = 0, x = 0, y = 0;
    for (int m = 0; m < P; ++m) {
        const double w = int_tri[(size_t)m * N + i];
        t += ts[(size_t)m * N + i] * w;
        x += qxs[(size_t)m * N + i] * w;
        y += qys[(size_t)m * N + i] * w;
    }
    temp[i] = t;
    qx[i] = x;
    qy[i] = y;
}

// [rows][P] binary64 -> [rows][pad4(P)] in T
template <typename T>
__global__ void __launch_bounds__(256) pad_rows_kernel(const double* in, T* out, int rows, int P)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (int64_t)rows * P) return;
    const int64_t r = g / P;
    const int m = (int)(g - r * P);
    out[r * pad4(P) + m] = (T)in[g];
}

inline unsigned nblocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

template <typename F>
int guarded(pbte_ctx* ctx, F&& f)
{
    if (!ctx) return -1;
    try {
        return f();
    } catch (const std::string& e) {
        ctx->err = e;
        return -2;
    }
}

}  // namespace

// ==========================================================================
// C-ABI
// ==========================================================================
namespace {
// Kr[a][q] = integral over the unit edge of l_a(s) l_{deg-q}(s), l_k the
// Lagrange polynomials on the equispaced edge nodes k / deg: the element
// across lists the shared edge's nodes in the opposite order
std::vector<double> reversed_edge_matrix(int deg)
{
    const int Q = deg + 1;
    std::vector<std::vector<double>> c(Q, std::vector<double>(Q, 0.0));    // coefficients of l_k
    for (int k = 0; k < Q; ++k) {
        std::vector<double> p{1.0};
        for (int j = 0; j < Q; ++j) {
            if (j == k) continue;
            const double sj = (double)j / deg, den = (double)(k - j) / deg;
            std::vector<double> np(p.size() + 1, 0.0);
            for (size_t e = 0; e < p.size(); ++e) {
                np[e + 1] += p[e] / den;
                np[e] -= p[e] * sj / den;
            }
            p = np;
        }
        c[k] = p;
    }
    std::vector<double> kr((size_t)Q * Q, 0.0);
    for (int a = 0; a < Q; ++a)
        for (int q = 0; q < Q; ++q) {
            const auto& x = c[a];
            const auto& y = c[deg - q];
            double v = 0.0;
            for (size_t e = 0; e < x.size(); ++e)
                for (size_t f = 0; f < y.size(); ++f) v += x[e] * y[f] / (double)(e + f + 1);
            kr[(size_t)a * Q + q] = v;
        }
    return kr;
}

// workgroups of `block` threads of the binary32 pipelined sweep kernel that
// one CU holds at once by their registers (LDS aside); 0 when unknown
template <int P>
int pipe_blocks_per_cu(int depth, bool one_dir, int block)
{
    constexpr int pf = lu_stored_mode<P, float>();
    const void* fn =
        depth == 2 ? (one_dir ? reinterpret_cast<const void*>(sweep_pipe_kernel<P, float, pf, false, 2, 1>)
                              : reinterpret_cast<const void*>(sweep_pipe_kernel<P, float, pf, false, 2, 0>))
                   : (one_dir ? reinterpret_cast<const void*>(sweep_pipe_kernel<P, float, pf, false, 3, 1>)
                              : reinterpret_cast<const void*>(sweep_pipe_kernel<P, float, pf, false, 3, 0>));
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, block, 0) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

}  // namespace

extern "C" const char* pbte_last_error(pbte_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

extern "C" void pbte_ctx_destroy(pbte_ctx* ctx) { delete ctx; }

extern "C" pbte_ctx* pbte_ctx_create(const pbte_dims* dims, const pbte_setup* su, void* stream,
                                     char* err, int32_t errlen)
{
    auto fail = [&](const std::string& m) -> pbte_ctx* {
        if (err && errlen > 0) {
            std::strncpy(err, m.c_str(), (size_t)errlen - 1);
            err[errlen - 1] = 0;
        }
        return nullptr;
    };
    if (!dims || !su) return fail("null dims/setup");
    const int P = dims->nd, Q = dims->nf, N = dims->n_tris, D = dims->ndir, F = dims->n_faces;
    if (!((P == 3 && Q == 2) || (P == 6 && Q == 3) || (P == 10 && Q == 4) || (P == 15 && Q == 5)))
        return fail("unsupported (ndof_tri, ndof_fc): degrees 1-4 only");
    if (N <= 0 || D <= 0) return fail("empty mesh or velocity set");
    pbte_ctx* ctx = new pbte_ctx();
    // PBTE_SETUP_TIMES set: the wall time of each phase on stderr
    const bool say = std::getenv("PBTE_SETUP_TIMES") != nullptr;
    auto t_last = std::chrono::steady_clock::now();
    auto phase = [&](const char* name) {
        if (!say) return;
        const auto now = std::chrono::steady_clock::now();
        std::fprintf(stderr, "pbte_ctx_create: %-24s %8.3f s\n", name,
                     std::chrono::duration<double>(now - t_last).count());
        t_last = now;
    };
    try {
        ctx->dm = *dims;
        ctx->P = P; ctx->Q = Q; ctx->N = N; ctx->D = D; ctx->F = F; ctx->deg = Q - 1;
        ctx->stream = (hipStream_t)stream;
        ctx->cv = su->cv; ctx->vg = su->vg; ctx->tau_r = su->tau_r; ctx->tau_n = su->tau_n;
        ctx->tau_c = su->tau_c; ctx->pi = su->pi;
        // items per workgroup and level: at most 256 (a binary32 pipelined
        // workgroup of 512 items, which its registers allow, took 1.85x
        // longer per level for 1.5x the items on the 256^2 P2 case)
        const int max_items = 256;
        ctx->block = su->block > 0 ? su->block : 256;
        if (ctx->block % 64 || ctx->block > max_items)
            throw std::string("block must be a multiple of 64, <= ") + std::to_string(max_items);
        ctx->variant = su->sweep_variant;
        if (ctx->variant < 0 || ctx->variant > 4) throw std::string("sweep_variant must be 0-4");
        ctx->items = ctx->block;
        const int deg = Q - 1;

        // --- direction tables -------------------------------------------------
        ctx->cxv = ctx->upload(su->cxv, D);
        ctx->cyv = ctx->upload(su->cyv, D);
        ctx->domega = ctx->upload(su->domega, D);
        std::vector<float> f32(D);
        for (int d = 0; d < D; ++d) f32[d] = (float)su->cxv[d];
        ctx->cxv32 = ctx->upload(f32.data(), D);
        for (int d = 0; d < D; ++d) f32[d] = (float)su->cyv[d];
        ctx->cyv32 = ctx->upload(f32.data(), D);

        ctx->mass = ctx->upload(su->mass, (size_t)P * P * N);
        ctx->gx = ctx->upload(su->gx, (size_t)N * P * P);
        ctx->gy = ctx->upload(su->gy, (size_t)N * P * P);
        ctx->fcm = ctx->upload(su->fcm, (size_t)N * 3 * P * P);
        ctx->int_tri = ctx->upload(su->int_tri, (size_t)P * N);
        {
            std::vector<double> nrm((size_t)N * 6);
            for (int i = 0; i < N; ++i)
                for (int il = 0; il < 3; ++il) {
                    nrm[(size_t)i * 6 + 2 * il] = su->nx[i * 3 + il];
                    nrm[(size_t)i * 6 + 2 * il + 1] = su->ny[i * 3 + il];
                }
            ctx->nrm = ctx->upload(nrm.data(), nrm.size());
        }

        phase("direction tables");
        // --- element data -----------------------------------------------------
        // CellGeo (factorisation), the compact face blocks and thermal wall
        // vectors (transport residual), the face lengths and the elements
        // across
        std::vector<CellGeo<double>> g64(N);
        const int KQ = pad4(Q * Q);
        std::vector<double> K((size_t)N * 3 * KQ, 0.0);
        std::vector<double> bvec((size_t)std::max(F, 1) * pad4(P) + 32, 0.0);
        std::vector<double> flen((size_t)N * 3, 0.0);
        std::vector<int32_t> h_other((size_t)N * 3, -1);
        // Kr from the first interior or periodic face block over its length:
        // the reference integrates every face with the same edge rule, so the
        // blocks are proportional to rounding (the exact edge matrix differs
        // from them by that rule's own error, ~1e-13); the exact matrix only
        // when there is no such face
        std::vector<double> kr = reversed_edge_matrix(deg);
        bool have_kr = false;
        double worst = 0.0, kworst = 0.0;
        for (int i = 0; i < N; ++i) {
            for (int il = 0; il < 3; ++il) {
                g64[i].n[2 * il] = su->nx[i * 3 + il];
                g64[i].n[2 * il + 1] = su->ny[i * 3 + il];
                const int fc = su->tri_faces[i * 3 + il];
                const int bc = su->face_bc[fc];
                const double* tt = su->ttfc + ((size_t)i * 3 + il) * P * P;
                int code, other, ilp = 0;
                if (bc < 0) {
                    code = FACE_INTERIOR;
                    other = su->neighbour[i * 3 + il];
                    ilp = su->other_lfc[i * 3 + il];
                } else {
                    const int typ = su->bc_type[bc];
                    if (typ == 3) {
                        code = FACE_PERIODIC;
                        other = su->periodic_tri[fc];
                        ilp = su->other_lfc[i * 3 + il];
                    } else if (typ == 2) {
                        code = FACE_ADIABATIC;
                        other = fc;
                    } else if (typ == 1) {
                        code = FACE_THERMAL;
                        other = fc;
                        const double tw = su->cv / 4.0 / su->pi * su->bc_temp[bc];
                        for (int m = 0; m < P; ++m) bvec[(size_t)fc * pad4(P) + m] = tw * tt[m * P + 0];
                    } else {
                        throw std::string("boundary type not supported by the sweep");
                    }
                }
                std::vector<char> onr(P, 0);
                for (int a = 0; a < Q; ++a) onr[face_node(deg, il, a)] = 1;
                if (code <= FACE_PERIODIC) {
                    if (other < 0 || ilp < 0 || ilp > 2) throw std::string("bad across-face element");
                    h_other[(size_t)i * 3 + il] = other;
                    double bmax = 0.0, omax = 0.0, len = 0.0;
                    std::vector<char> onc(P, 0);
                    for (int a = 0; a < Q; ++a) onc[face_node(deg, ilp, a)] = 1;
                    for (int m = 0; m < P; ++m)
                        for (int l = 0; l < P; ++l) {
                            const double v = std::fabs(tt[m * P + l]);
                            if (onr[m] && onc[l]) bmax = std::max(bmax, v);
                            else omax = std::max(omax, v);
                        }
                    if (bmax > 0) worst = std::max(worst, omax / bmax);
                    for (int a = 0; a < Q; ++a)
                        for (int q = 0; q < Q; ++q) {
                            const double v = tt[face_node(deg, il, a) * P + face_node(deg, ilp, q)];
                            K[((size_t)i * 3 + il) * KQ + a * Q + q] = v;
                            len += v;
                        }
                    if (!have_kr && len > 0.0) {
                        for (int k = 0; k < Q * Q; ++k)
                            kr[k] = K[((size_t)i * 3 + il) * KQ + k] / len;
                        have_kr = true;
                    }
                    flen[(size_t)i * 3 + il] = len;
                } else {
                    // a wall face: int_tri_tri_fc holds the integral of each
                    // basis function on the face in every column
                    double len = 0.0, bmax = 0.0, omax = 0.0;
                    for (int a = 0; a < Q; ++a) len += tt[face_node(deg, il, a) * P];
                    for (int m = 0; m < P; ++m) {
                        const double v = std::fabs(tt[m * P]);
                        if (onr[m]) bmax = std::max(bmax, v); else omax = std::max(omax, v);
                    }
                    if (bmax > 0) worst = std::max(worst, omax / bmax);
                    // the adiabatic emission goes through the face mass matrix
                    const double* fm = su->fcm + ((size_t)i * 3 + il) * P * P;
                    double fmax = 0.0, fout = 0.0;
                    for (int m = 0; m < P; ++m)
                        for (int l = 0; l < P; ++l) {
                            const double v = std::fabs(fm[m * P + l]);
                            if (onr[m] && onr[l]) fmax = std::max(fmax, v); else fout = std::max(fout, v);
                        }
                    if (fmax > 0) worst = std::max(worst, fout / fmax);
                    flen[(size_t)i * 3 + il] = len;
                }
                g64[i].other[il] = other;
                g64[i].info[il] = code | (ilp << 2) | (fc << 4);
            }
        }
        if (worst > 1e-9)
            throw std::string("element x neighbour face tensor is not confined to the shared face "
                              "(non-conforming mesh?)");
        // every face block must be its length times Kr; a wall face's column
        // its length times the row sums of Kr
        for (int i = 0; i < N; ++i)
            for (int il = 0; il < 3; ++il) {
                const double len = flen[(size_t)i * 3 + il];
                const int code = g64[i].info[il] & 3;
                if (code <= FACE_PERIODIC) {
                    const double* Kf = K.data() + ((size_t)i * 3 + il) * KQ;
                    double bmax = 0.0, e = 0.0;
                    for (int k = 0; k < Q * Q; ++k) {
                        bmax = std::max(bmax, std::fabs(Kf[k]));
                        e = std::max(e, std::fabs(Kf[k] - len * kr[k]));
                    }
                    if (bmax > 0) kworst = std::max(kworst, e / bmax);
                } else {
                    const double* tt = su->ttfc + ((size_t)i * 3 + il) * P * P;
                    double bmax = 0.0, e = 0.0;
                    for (int a = 0; a < Q; ++a) {
                        double rsum = 0.0;
                        for (int q = 0; q < Q; ++q) rsum += kr[a * Q + q];
                        const double v = tt[face_node(deg, il, a) * P];
                        bmax = std::max(bmax, std::fabs(v));
                        e = std::max(e, std::fabs(v - len * rsum));
                    }
                    if (bmax > 0) kworst = std::max(kworst, e / bmax);
                }
            }
        if (kworst > 1e-11)
            throw std::string("face blocks are not a length times the unit-edge mass matrix "
                              "(curved or non-affine faces?)");
        {
            // the fused sweep's outflow term: every face mass block
            // int_tri_fc[i, il] is its length times one unit-edge mass
            // matrix Kf on the face's nodes (own node order) and zero
            // elsewhere; otherwise that sweep stays off
            std::vector<double> kf((size_t)Q * Q, 0.0);
            bool ok = N > 0 && flen[0] > 0.0;
            if (ok)
                for (int r = 0; r < Q; ++r)
                    for (int q = 0; q < Q; ++q)
                        kf[(size_t)r * Q + q] = su->fcm[(size_t)face_node(deg, 0, r) * P + face_node(deg, 0, q)] / flen[0];
            for (int i = 0; i < N && ok; ++i)
                for (int il = 0; il < 3 && ok; ++il) {
                    const double* fm = su->fcm + ((size_t)i * 3 + il) * P * P;
                    const double len = flen[(size_t)i * 3 + il];
                    double bmax = 0.0, e = 0.0;
                    std::vector<int> on(P, 0);
                    for (int r = 0; r < Q; ++r) on[face_node(deg, il, r)] = 1 + r;
                    for (int m = 0; m < P; ++m)
                        for (int l = 0; l < P; ++l) {
                            const double v = fm[m * P + l];
                            const double want = on[m] && on[l] ? len * kf[(size_t)(on[m] - 1) * Q + on[l] - 1] : 0.0;
                            bmax = std::max(bmax, std::fabs(v));
                            e = std::max(e, std::fabs(v - want));
                        }
                    ok = bmax > 0.0 && e <= 1e-11 * bmax;
                }
            ctx->fused_ok = ok;
            ctx->h_fref.assign((size_t)3 * P * P + (size_t)Q * Q, 0.0);
            for (int k = 0; k < Q * Q; ++k) ctx->h_fref[(size_t)3 * P * P + k] = kf[k];
        }
        ctx->geo64 = ctx->upload(g64.data(), N);
        ctx->h_flen.assign(flen.begin(), flen.end());
        ctx->K64 = ctx->upload(K.data(), K.size());
        ctx->bvec64 = ctx->upload(bvec.data(), bvec.size());
        ctx->other = ctx->upload(h_other.data(), h_other.size());
        for (int k = 0; k < Q * Q; ++k) {
            ctx->kref64[k] = kr[k];
            ctx->kref32[k] = (float)kr[k];
        }
        {
            // cell records: source (written by every sweep), length-scaled
            // outward normals, info words (bit patterns)
            const int rs64 = rec_stride_rt(P, 8), rs32 = rec_stride_rt(P, 4);
            const int RN = 3 * P, RI = 3 * P + 6;
            std::vector<double> c64((size_t)N * rs64, 0.0);
            std::vector<float> c32((size_t)N * rs32, 0.0f);
            for (int i = 0; i < N; ++i) {
                double* r = c64.data() + (size_t)i * rs64;
                float* r32 = c32.data() + (size_t)i * rs32;
                for (int il = 0; il < 3; ++il) {
                    const double len = flen[(size_t)i * 3 + il];
                    r[RN + 2 * il] = g64[i].n[2 * il] * len;
                    r[RN + 2 * il + 1] = g64[i].n[2 * il + 1] * len;
                    r32[RN + 2 * il] = (float)r[RN + 2 * il];
                    r32[RN + 2 * il + 1] = (float)r[RN + 2 * il + 1];
                    const uint32_t bits = (uint32_t)g64[i].info[il];
                    const uint64_t b64 = bits;
                    std::memcpy(r + RI + il, &b64, sizeof(double));
                    std::memcpy(r32 + RI + il, &bits, sizeof(float));
                }
            }
            ctx->rs64 = rs64;
            ctx->rs32 = rs32;
            ctx->crec64 = ctx->upload(c64.data(), c64.size());
            ctx->crec32 = ctx->upload(c32.data(), c32.size());
        }
        // wall vectors over the face length, one row per thermal or
        // adiabatic wall face: -tw ttfc[:, 0] / L for a thermal wall (the
        // inflow term of _kernels.py:213-215), flux_wall / L for an adiabatic
        // one, refreshed by every sweep (_kernels.py:210-212)
        const int TQ = trace_pad(Q);
        const int PV = pad4(P);
        std::vector<int32_t> wrow((size_t)std::max(F, 1), -1);
        {
            std::vector<double> w64;
            std::vector<int32_t> ar, af;
            std::vector<double> arl;
            for (int f = 0; f < F; ++f) {
                const int bc = su->face_bc[f];
                if (bc < 0) continue;
                const int typ = su->bc_type[bc];
                if (typ != 1 && typ != 2) continue;
                int tri = su->face_tri[f * 2], lfc = su->face_lfc[f * 2];
                if (tri < 0) { tri = su->face_tri[f * 2 + 1]; lfc = su->face_lfc[f * 2 + 1]; }
                wrow[f] = ctx->n_wrow++;
                const double len = flen[(size_t)tri * 3 + lfc];
                const double tw = typ == 1 ? su->cv / 4.0 / su->pi * su->bc_temp[bc] : 0.0;
                const double* tt = su->ttfc + ((size_t)tri * 3 + lfc) * P * P;
                for (int m = 0; m < PV; ++m) w64.push_back(typ == 1 && m < P ? -(tw * tt[m * P]) / len : 0.0);
                if (typ == 2) {
                    ar.push_back(wrow[f]); af.push_back(f);
                    arl.push_back(1.0 / len);
                }
            }
            w64.resize(std::max<size_t>(w64.size(), 1), 0.0);
            std::vector<float> w32(w64.begin(), w64.end());
            ctx->wtr64 = ctx->upload(w64.data(), w64.size());
            ctx->wtr32 = ctx->upload(w32.data(), w32.size());
            ctx->n_adw = (int)ar.size();
            ctx->adw_row = ctx->upload(ar.data(), ar.size());
            ctx->adw_face = ctx->upload(af.data(), af.size());
            ctx->adw_rlen = ctx->upload(arl.data(), arl.size());
        }
        if (ctx->n_wrow >= kWallThermal) throw std::string("too many wall faces");

        // --- adiabatic walls --------------------------------------------------
        std::vector<int32_t> wf, wt, wl;
        for (int f = 0; f < F; ++f) {
            const int bc = su->face_bc[f];
            if (bc < 0 || su->bc_type[bc] != 2) continue;
            int tri = su->face_tri[f * 2], lfc = su->face_lfc[f * 2];
            if (tri < 0) { tri = su->face_tri[f * 2 + 1]; lfc = su->face_lfc[f * 2 + 1]; }
            wf.push_back(f); wt.push_back(tri); wl.push_back(lfc);
        }
        ctx->n_wall = (int)wf.size();
        ctx->wall_face = ctx->upload(wf.data(), wf.size());
        ctx->wall_tri = ctx->upload(wt.data(), wt.size());
        ctx->wall_lfc = ctx->upload(wl.data(), wl.size());
        std::vector<int32_t> pf, pt;
        for (int f = 0; f < F; ++f)
            if (su->periodic_tri[f] >= 0) { pf.push_back(f); pt.push_back(su->periodic_tri[f]); }
        ctx->n_per = (int)pf.size();
        ctx->per_face = ctx->upload(pf.data(), pf.size());
        ctx->per_tri = ctx->upload(pt.data(), pt.size());
        if (N > kCellMask || F > kFaceValMask) throw std::string("mesh too large for the record fields");

        phase("element data");
        // --- bundles: directions with identical inflow signs ------------------
        std::vector<uint64_t> hsh(D);
        std::vector<std::vector<uint8_t>> sig(D);
        {
            // directions split over host threads, each writing its own rows
            std::atomic_int next{0};
            auto signs = [&]() {
                for (int d = next++; d < D; d = next++) {
                    auto& sg = sig[d];
                    sg.resize((size_t)N * 3);
                    uint64_t h = 1469598103934665603ull;
                    for (int i = 0; i < N; ++i)
                        for (int il = 0; il < 3; ++il) {
                            const double sp = su->cxv[d] * su->nx[i * 3 + il] + su->cyv[d] * su->ny[i * 3 + il];
                            const uint8_t v = sp < 0.0 ? 1 : (sp > 0.0 ? 2 : 0);
                            sg[(size_t)i * 3 + il] = v;
                            h = (h ^ v) * 1099511628211ull;
                        }
                    hsh[d] = h;
                }
            };
            const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
            std::vector<std::thread> th;
            for (unsigned t = 0; t < hw; ++t) th.emplace_back(signs);
            for (auto& x : th) x.join();
        }
        int bmax = su->max_bundle > 0 ? su->max_bundle : 1;
        bmax = std::min(bmax, 8);
        while (ctx->items % bmax) --bmax;
        std::vector<std::vector<int>> groups;
        {
            std::vector<char> used(D, 0);
            for (int d = 0; d < D; ++d) {
                if (used[d]) continue;
                std::vector<int> grp{d};
                used[d] = 1;
                for (int e = d + 1; e < D; ++e)
                    if (!used[e] && hsh[e] == hsh[d] && sig[e] == sig[d] &&
                        std::memcmp(su->order + (size_t)e * N, su->order + (size_t)d * N,
                                    (size_t)N * sizeof(int32_t)) == 0) {
                        grp.push_back(e);
                        used[e] = 1;
                    }
                groups.push_back(grp);
            }
        }
        std::vector<std::vector<int>> bundles;
        std::vector<int> bundle_group;
        for (size_t gi = 0; gi < groups.size(); ++gi) {
            const auto& grp = groups[gi];
            size_t k = 0;
            while (k < grp.size()) {
                int take = bmax;
                while (take > 1 && (size_t)take > grp.size() - k) take >>= 1;
                bundles.emplace_back(grp.begin() + k, grp.begin() + k + take);
                bundle_group.push_back((int)gi);
                k += take;
            }
        }
        const int nb = (int)bundles.size();
        ctx->nb = nb;
        int bm = 1;
        for (auto& bb : bundles) bm = std::max(bm, (int)bb.size());
        phase("bundles");
        // inflow wall elements per bundle (the same for every direction of it)
        {
            int most = 0;
            std::vector<int32_t> mark((size_t)N, -1);
            for (int k = 0; k < nb; ++k) {
                const int d = bundles[k][0];
                int n = 0;
                for (int f = 0; f < F; ++f) {
                    if (wrow[f] < 0) continue;
                    int tri = su->face_tri[f * 2], lfc = su->face_lfc[f * 2];
                    if (tri < 0) { tri = su->face_tri[f * 2 + 1]; lfc = su->face_lfc[f * 2 + 1]; }
                    const double sp = su->cxv[d] * su->nx[tri * 3 + lfc] + su->cyv[d] * su->ny[tri * 3 + lfc];
                    if (sp < 0.0 && mark[tri] != k) { mark[tri] = k; ++n; }
                }
                most = std::max(most, n);
            }
            ctx->max_wall_in = most;
        }
        // LDS pool: what one workgroup may hold when every CU runs its share
        // of the bundles at once, counted at binary32.  For binary64 sweeps
        // (lds_value_bytes 8) half of a workgroup's LDS: a binary64 sweep
        // takes twice the bytes (and by its register use runs one workgroup
        // of up to 256 threads per CU in any case).  For binary32 sweeps all
        // of it; a binary64 sweep whose pool then does not fit runs the
        // plain level kernel.
        {
            int dev = 0, ncu = 1, lds_blk = 65536, lds_cu = 0;
            HIP_TRY(hipGetDevice(&dev));
            HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
            HIP_TRY(hipDeviceGetAttribute(&lds_blk, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
            if (hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) !=
                    hipSuccess || lds_cu <= 0)
                lds_cu = lds_blk;
            ctx->depth = ctx->variant >= 2 ? ctx->variant : 3;
            // bundles a CU runs at once: its share of them, but no more
            // workgroups than its registers hold (the widest workgroup)
            int per_cu = std::max(1, (nb + std::max(ncu, 1) - 1) / std::max(ncu, 1));
            int held = 0;
            switch (P) {
            case 3: held = pipe_blocks_per_cu<3>(ctx->depth, bm == 1, 256); break;
            case 6: held = pipe_blocks_per_cu<6>(ctx->depth, bm == 1, 256); break;
            case 10: held = pipe_blocks_per_cu<10>(ctx->depth, bm == 1, 256); break;
            case 15: held = pipe_blocks_per_cu<15>(ctx->depth, bm == 1, 256); break;
            }
            if (held > 0) per_cu = std::min(per_cu, held);
            const size_t blk = su->lds_value_bytes == 4 ? (size_t)lds_blk : (size_t)lds_blk / 2;
            const size_t budget = std::min<size_t>(blk, (size_t)lds_cu / per_cu);
            ctx->lds_blk = (size_t)lds_blk;
            const size_t slot_bytes = (size_t)bm * pad4(3 * TQ) * 4;
            const size_t wall_bytes = (size_t)ctx->max_wall_in * bm * PV * 4;
            if (budget < wall_bytes + 4 * slot_bytes) {
                // the bundles' wall inflow does not fit beside a pool: the
                // plain level kernel (wall rows and upwind values read from
                // global memory) sweeps this context
                ctx->variant = 1;
                ctx->pool_slots = 1;
            } else {
                int pool = std::min((int)((budget - wall_bytes) / slot_bytes) - 2, kPoolMax);
                if (su->lds_slots_cap > 0) pool = std::min(pool, su->lds_slots_cap);
                ctx->pool_slots = std::max(pool, 1);
            }
        }
        phase("wall lists, LDS budget");
        // Schedule classes: the bundles of one direction group (the same
        // order and the same inflow sign on every face) and one size have
        // the same levels, slots, face records, wall lists and pool; each
        // class is built once and stored once, and its bundles point at it
        // (the records of a class are then read by all of its bundles, which
        // the block map below places on one XCD, out of that XCD's L2).
        std::vector<int> cls_of(nb), rep;
        {
            std::map<std::pair<int, int>, int> id;
            for (int k = 0; k < nb; ++k) {
                const auto key = std::make_pair(bundle_group[k], (int)bundles[k].size());
                auto it = id.find(key);
                if (it == id.end()) {
                    it = id.emplace(key, (int)rep.size()).first;
                    rep.push_back(k);
                }
                cls_of[k] = it->second;
            }
        }
        const int nc = (int)rep.size();
        if (say) std::fprintf(stderr, "pbte_ctx_create: %d direction groups, %d bundles, %d schedule classes\n",
                              (int)groups.size(), nb, nc);
        std::vector<int32_t> h_slot((size_t)nc * N), h_lev_ptr;
        std::vector<int32_t> h_ndir(nb), h_dir_off(nb + 1, 0), h_dirs;
        std::vector<int64_t> h_lu_off(nb + 1, 0), h_item_off(nb + 1, 0);
        std::vector<std::vector<int32_t>> slots(nc);
        std::vector<std::vector<int32_t>> lptr(nc);
        std::vector<std::vector<int4>> wlists(nc);
        std::vector<int4> h_rec((size_t)nc * N);
        std::vector<uint8_t> global_flag(nc, 0);
        std::vector<int64_t> peak(nc, 0), slots_used(nc, 0);
        // per class: reads by kind (previous-sweep non-periodic, previous-
        // sweep periodic, periodic), whether its staged reads are encoded as
        // such, and its stage list with the lengths of its three runs
        std::vector<std::array<int64_t, 3>> kind_reads(nc, {0, 0, 0});
        std::vector<uint8_t> staged(nc, 0);
        std::vector<std::vector<int32_t>> stage_lists(nc);
        std::vector<std::array<int32_t, 3>> stage_cnt(nc, {0, 0, 0});
        // levels of every class: slots in level order (the reference order
        // within a level) and the level boundaries
        std::vector<std::vector<int32_t>> lcnt(nc);
        auto run_threads = [&](auto&& fn) {
            unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
            std::vector<std::thread> th;
            for (unsigned t = 0; t < hw; ++t) th.emplace_back(fn);
            for (auto& x : th) x.join();
        };
        {
            std::atomic_int next{0};
            run_threads([&]() {
                std::vector<int32_t> level_of;
                for (int k = next++; k < nc; k = next++) {
                    const int d = bundles[rep[k]][0];
                    const int32_t* order = su->order + (size_t)d * N;
                    int nlev = pbte_detail_levels(N, order, su->neighbour, su->tri_faces, su->face_bc,
                                                  su->bc_type, su->periodic_tri, su->nx, su->ny, su->cxv[d],
                                                  su->cyv[d], level_of);
                    std::vector<int32_t> cnt(nlev + 1, 0);
                    for (int i = 0; i < N; ++i) ++cnt[level_of[i] + 1];
                    for (int l = 0; l < nlev; ++l) cnt[l + 1] += cnt[l];
                    std::vector<int32_t> start(cnt.begin(), cnt.end() - 1);
                    auto& sl = slots[k];
                    sl.assign(N, -1);
                    for (int p = 0; p < N; ++p) {
                        const int i = order[p];
                        sl[start[level_of[i]]++] = i;
                    }
                    lcnt[k] = std::move(cnt);
                }
            });
        }
        phase("levels");
        if (ctx->variant != 1 && su->block <= 0) {
            // the pipelined kernel runs one wave per 64 items of the widest
            // level (every wave issues the full load stream of a level), at
            // most three waves: a level's time grows with the waves that run
            // it, steeply from three to four (256 x 256 P2, 512 ordinates,
            // paired directions, one MI355X: 1.04 / 1.19 / 1.53 us per level
            // at two / three / four waves; the sweep 2.69 / 2.33 / 2.44 ms
            // with the sublevels that the narrower width adds)
            constexpr int kAutoItems = 192;
            int64_t widest = 1;
            for (int k = 0; k < nc; ++k)
                for (size_t l = 0; l + 1 < lcnt[k].size(); ++l)
                    widest = std::max<int64_t>(widest,
                                               (int64_t)(lcnt[k][l + 1] - lcnt[k][l]) * bundles[rep[k]].size());
            ctx->block = ctx->items = (int)std::min<int64_t>(kAutoItems, (widest + 63) / 64 * 64);
            for (auto& bb : bundles)
                if (ctx->items % (int)bb.size()) ctx->block = ctx->items = kAutoItems;
        }
        std::atomic_int next{0};
        // a paired schedule asked to fail on overflow (pairs_or_fail) stops
        // at the first bundle whose pool overflows: the caller builds the
        // one-direction schedule instead
        std::atomic<bool> pair_overflow{false};
        struct Read { int code, val, ilp; };
        auto work = [&]() {
            // per-thread work arrays, reused by every class the thread builds
            // (a fresh allocation per class page-faults tens of MB per thread)
            std::vector<int32_t> sub_of((size_t)N), pos((size_t)N), slot_of((size_t)N), last_reader((size_t)N);
            std::vector<int32_t> sid((size_t)N);
            std::vector<Read> rd((size_t)N * 3);
            for (int k = next++; k < nc; k = next++) {
                if (pair_overflow.load(std::memory_order_relaxed)) break;
                const int d = bundles[rep[k]][0];
                const int32_t* order = su->order + (size_t)d * N;
                const auto& cnt = lcnt[k];
                const int nlev = (int)cnt.size() - 1;
                auto& sl = slots[k];
                const int B = (int)bundles[rep[k]].size();
                const int width = ctx->items / B;
                auto& lp = lptr[k];
                lp.clear();
                for (int l = 0; l < nlev; ++l)
                    for (int s = cnt[l]; s < cnt[l + 1]; s += width) lp.push_back(s);
                lp.push_back(N);
                // sublevel of every slot
                for (size_t j = 0; j + 1 < lp.size(); ++j)
                    for (int q = lp[j]; q < lp[j + 1]; ++q) sub_of[q] = (int32_t)j;
                // per-slot records: the element and, per local face, what the
                // sweep reads through it (REC_*).  This sweep's upwind values
                // come from the LDS pool slot of the upwind element; a bundle
                // whose pool would overflow reads them from global memory.
                for (int p = 0; p < N; ++p) pos[order[p]] = p;
                for (int q = 0; q < N; ++q) slot_of[sl[q]] = q;
                const double cxd = su->cxv[d], cyd = su->cyv[d];
                std::fill(last_reader.begin(), last_reader.end(), -1);
                auto& wlst = wlists[k];
                wlst.clear();
                for (int q = 0; q < N; ++q) {
                    const int i = sl[q];
                    int wk = -1;
                    for (int il = 0; il < 3; ++il) {
                        Read r{REC_NONE, 0, 0};
                        const double sp = cxd * su->nx[i * 3 + il] + cyd * su->ny[i * 3 + il];
                        if (sp < 0.0) {
                            const int fc = su->tri_faces[i * 3 + il];
                            const int bc = su->face_bc[fc];
                            const int typ = bc < 0 ? 0 : su->bc_type[bc];
                            if (bc < 0 || typ == 3) {
                                const int j = bc < 0 ? su->neighbour[i * 3 + il] : su->periodic_tri[fc];
                                const bool per = bc >= 0;
                                const int ilp = su->other_lfc[i * 3 + il];
                                if (pos[j] < pos[i]) {
                                    const int u = slot_of[j];
                                    r = {per ? REC_PER_LDS : REC_UP_LDS, u, ilp};
                                    last_reader[u] = std::max(last_reader[u], sub_of[q]);
                                } else {
                                    r = {per ? REC_PER_OLD : REC_OLD, j, ilp};
                                }
                            } else {
                                if (wk < 0) {
                                    wk = (int)wlst.size();
                                    wlst.push_back(make_int4(i, -1, -1, -1));
                                }
                                const int ent = wrow[fc] | (typ == 1 ? kWallThermal : 0);
                                int4& e = wlst[wk];
                                (il == 0 ? e.y : (il == 1 ? e.z : e.w)) = ent;
                                r = {typ == 2 ? REC_ADIABATIC : REC_THERMAL, wk, 0};
                            }
                        }
                        rd[(size_t)q * 3 + il] = r;
                    }
                }
                // pool slots: interval allocation in slot order; a slot is free
                // again once the sublevel of its last reader has passed
                std::fill(sid.begin(), sid.end(), -1);
                bool overflow = false;
                {
                    using Busy = std::pair<int32_t, int32_t>;     // (last reader sublevel, slot)
                    std::priority_queue<Busy, std::vector<Busy>, std::greater<Busy>> busy;
                    std::vector<int32_t> freelist;
                    int used = 0;
                    for (int u = 0; u < N && !overflow; ++u) {
                        if (last_reader[u] < 0) continue;
                        while (!busy.empty() && busy.top().first < sub_of[u]) {
                            freelist.push_back(busy.top().second);
                            busy.pop();
                        }
                        int id;
                        if (!freelist.empty()) { id = freelist.back(); freelist.pop_back(); }
                        else if (used < ctx->pool_slots) id = used++;
                        else { overflow = true; break; }
                        sid[u] = id;
                        busy.push({last_reader[u], id});
                        peak[k] = std::max<int64_t>(peak[k], (int64_t)busy.size());
                    }
                    slots_used[k] = overflow ? 0 : used;
                }
                global_flag[k] = overflow ? 1 : 0;
                if (overflow && su->pairs_or_fail && bundles[rep[k]].size() > 1) {
                    pair_overflow = true;
                    break;
                }
                // stage lists (a bundle with a pool only): one entry per
                // staged read, (element or periodic face) << 2 | face across,
                // in three runs: previous-sweep reads across non-periodic
                // faces, across periodic faces, every periodic read
                {
                    int64_t n_np = 0, n_po = 0, n_per = 0;
                    for (size_t f = 0; f < (size_t)N * 3; ++f) {
                        n_np += rd[f].code == REC_OLD;
                        n_po += rd[f].code == REC_PER_OLD;
                        n_per += rd[f].code == REC_PER_OLD || rd[f].code == REC_PER_LDS;
                    }
                    if (!overflow && n_np + n_po + n_per > 0 && n_np + n_po < kStageMax && n_np + n_per < kStageMax) {
                        stage_cnt[k] = {(int32_t)n_np, (int32_t)n_po, (int32_t)n_per};
                        stage_lists[k].assign((size_t)(n_np + n_po + n_per), 0);
                        staged[k] = 1;
                    }
                }
                int32_t e_np = 0, e_po = 0, e_per = 0;
                for (int q = 0; q < N; ++q) {
                    int w[3];
                    for (int il = 0; il < 3; ++il) {
                        Read r = rd[(size_t)q * 3 + il];
                        if (r.code == REC_OLD) ++kind_reads[k][0];
                        if (r.code == REC_PER_OLD) ++kind_reads[k][1];
                        if (r.code == REC_PER_LDS || r.code == REC_PER_OLD) ++kind_reads[k][2];
                        if (r.code == REC_UP_LDS || r.code == REC_PER_LDS) {
                            if (overflow) r = {r.code == REC_UP_LDS ? REC_UP_GLOBAL : REC_PER_GLOBAL, sl[r.val], r.ilp};
                            else r.val = sid[r.val];
                        }
                        if (staged[k] && (r.code == REC_OLD || r.code == REC_PER_OLD || r.code == REC_PER_LDS)) {
                            // stage slots: without an explicit periodic
                            // buffer the first two runs, with one the
                            // first and the third
                            auto& lst = stage_lists[k];
                            const int n_np = stage_cnt[k][0], n_po = stage_cnt[k][1];
                            int hi = 0;
                            if (r.code != REC_OLD) {
                                const int fc = su->tri_faces[(size_t)sl[q] * 3 + il];
                                lst[(size_t)n_np + n_po + e_per] = fc << 2 | r.ilp;
                                hi = (n_np + e_per++) << kStageBits;
                            }
                            if (r.code == REC_OLD) {
                                lst[e_np] = r.val << 2 | r.ilp;
                                r = {REC_OLD_ST, e_np++, r.ilp};
                            } else if (r.code == REC_PER_OLD) {
                                lst[(size_t)n_np + e_po] = r.val << 2 | r.ilp;
                                r = {REC_PER_OLD_ST, (n_np + e_po++) | hi, r.ilp};
                            } else {
                                r = {REC_PER_LDS_ST, r.val | hi, r.ilp};
                            }
                        }
                        w[il] = face_word(r.code, r.ilp, r.val);
                    }
                    const int wst = overflow ? -1 : sid[q];
                    h_rec[(size_t)k * N + q] =
                        make_int4((int)((uint32_t)sl[q] | ((uint32_t)(wst + 1) << kCellBits)), w[0], w[1], w[2]);
                }
            }
        };
        run_threads(work);
        if (pair_overflow) throw std::string("paired bundles overflow the LDS trace pool (pairs_or_fail)");
        // the pool shrinks to the slots the bundles that kept it use (every
        // slot id is below that count; the zero and spill slots follow it)
        ctx->pool_slots = (int)std::max<int64_t>(1, *std::max_element(slots_used.begin(), slots_used.end()));
        int64_t maxw = 0;
        std::vector<int4> h_wcell;
        std::vector<int32_t> c_wc_off(nc + 1, 0), c_lev_off(nc, 0);
        for (int k = 0; k < nb; ++k) ctx->max_bundle_dirs = std::max(ctx->max_bundle_dirs, (int)bundles[k].size());
        for (int k = 0; k < nc; ++k) {
            const int B = (int)bundles[rep[k]].size();
            std::copy(slots[k].begin(), slots[k].end(), h_slot.begin() + (size_t)k * N);
            // lev_ptr is shared storage: each class's boundaries, its last
            // entry doubling as the end of its final level
            c_lev_off[k] = (int32_t)h_lev_ptr.size();
            for (size_t j = 0; j + 1 < lptr[k].size(); ++j) {
                h_lev_ptr.push_back(lptr[k][j]);
                maxw = std::max<int64_t>(maxw, (int64_t)(lptr[k][j + 1] - lptr[k][j]) * B);
            }
            h_lev_ptr.push_back(N);
            c_wc_off[k + 1] = c_wc_off[k] + (int32_t)wlists[k].size();
            h_wcell.insert(h_wcell.end(), wlists[k].begin(), wlists[k].end());
        }
        // per bundle: its directions, its factors, and where its class's
        // levels (lev_off: (start, end) into lev_ptr, lev_ptr[j+1] the end of
        // level j), records and slots (item_off: N of each per class) and
        // wall list (wc_off: (start, count)) are
        std::vector<int32_t> h_lev_se(2 * nb), h_wc_se(2 * nb);
        std::vector<uint8_t> b_global(nb);
        int64_t tot = 0;
        for (int k = 0; k < nb; ++k) {
            const int B = (int)bundles[k].size(), c = cls_of[k];
            h_ndir[k] = B;
            h_dir_off[k + 1] = h_dir_off[k] + B;
            for (int d : bundles[k]) h_dirs.push_back(d);
            h_lu_off[k + 1] = h_lu_off[k] + (int64_t)pad4(P * P) * N * B;
            h_item_off[k] = (int64_t)c * N;
            h_lev_se[2 * k] = c_lev_off[c];
            h_lev_se[2 * k + 1] = c_lev_off[c] + (int32_t)lptr[c].size() - 1;
            h_wc_se[2 * k] = c_wc_off[c];
            h_wc_se[2 * k + 1] = c_wc_off[c + 1] - c_wc_off[c];
            b_global[k] = global_flag[c];
            tot += (int64_t)lptr[c].size() - 1;
        }
        h_item_off[nb] = (int64_t)nc * N;
        ctx->total_levels = tot;
        ctx->max_width = maxw;
        ctx->n_classes = nc;
        phase("records, pools");
        // XCD-aware block -> bundle map: blocks are dispatched round robin over
        // 8 XCDs; give each XCD a contiguous range of bundles sorted by the
        // in-plane angle so its L2 sees similar sweep fronts.
        std::vector<int> sorted(nb);
        std::iota(sorted.begin(), sorted.end(), 0);
        std::stable_sort(sorted.begin(), sorted.end(), [&](int p, int q) {
            const int dp = bundles[p][0], dq = bundles[q][0];
            return std::atan2(su->cyv[dp], su->cxv[dp]) < std::atan2(su->cyv[dq], su->cxv[dq]);
        });
        std::vector<int32_t> h_block_bundle(nb);
        {
            const int X = 8;
            std::vector<int> count(X, 0), startx(X + 1, 0);
            for (int g = 0; g < nb; ++g) ++count[g % X];
            for (int x = 0; x < X; ++x) startx[x + 1] = startx[x] + count[x];
            for (int g = 0; g < nb; ++g) h_block_bundle[g] = sorted[startx[g % X] + g / X];
        }
        ctx->block_bundle = ctx->upload(h_block_bundle.data(), nb);
        ctx->bundle_ndir = ctx->upload(h_ndir.data(), nb);
        ctx->bundle_dir_off = ctx->upload(h_dir_off.data(), nb + 1);
        ctx->bundle_dirs = ctx->upload(h_dirs.data(), h_dirs.size());
        ctx->slot_cell = ctx->upload(h_slot.data(), h_slot.size());
        ctx->lev_off = ctx->upload(h_lev_se.data(), h_lev_se.size());
        ctx->lev_ptr = ctx->upload(h_lev_ptr.data(), h_lev_ptr.size());
        ctx->lu_off = ctx->upload(h_lu_off.data(), nb + 1);
        h_wcell.resize(std::max<size_t>(h_wcell.size(), 1), make_int4(0, -1, -1, -1));
        ctx->wcell = ctx->upload(h_wcell.data(), h_wcell.size());
        ctx->wc_off = ctx->upload(h_wc_se.data(), h_wc_se.size());
        phase("schedule upload");
        // per (class, slot): {cell | pool slot, face words}, read by every
        // direction of every bundle of the class
        ctx->irec = ctx->upload(h_rec.data(), h_rec.size());
        phase("item records");
        ctx->bundle_global = ctx->upload(b_global.data(), b_global.size());
        std::vector<int32_t> h_stage, c_stage_off(nc, 0);
        for (int c = 0; c < nc; ++c) {
            c_stage_off[c] = (int32_t)h_stage.size();
            h_stage.insert(h_stage.end(), stage_lists[c].begin(), stage_lists[c].end());
        }
        std::vector<int4> h_stage_off(nb);
        for (int k = 0; k < nb; ++k) {
            const int c = cls_of[k];
            ctx->global_bundles += global_flag[c];
            ctx->pool_peak = std::max(ctx->pool_peak, peak[c]);
            ctx->np_old_reads += kind_reads[c][0];
            ctx->per_old_reads += kind_reads[c][1];
            ctx->periodic_reads += kind_reads[c][2];
            if (!staged[c]) {
                ctx->np_old_unstaged += kind_reads[c][0];
                ctx->per_old_unstaged += kind_reads[c][1];
                ctx->per_unstaged += kind_reads[c][2];
            }
            const auto& sc = stage_cnt[c];
            h_stage_off[k] = make_int4(c_stage_off[c], sc[0], sc[1], sc[2]);
            ctx->stage_slots[0] = std::max(ctx->stage_slots[0], sc[0] + sc[1]);
            ctx->stage_slots[1] = std::max(ctx->stage_slots[1], sc[0] + sc[2]);
        }
        h_stage.resize(std::max<size_t>(h_stage.size(), 1), 0);
        ctx->stage_list = ctx->upload(h_stage.data(), h_stage.size());
        ctx->stage_off = ctx->upload(h_stage_off.data(), h_stage_off.size());
        ctx->item_off = ctx->upload(h_item_off.data(), nb + 1);
        ctx->h_lu_off = h_lu_off;
        ctx->lu_elems = h_lu_off[nb];
        ctx->item_elems = h_item_off[nb];

        // --- scratch -----------------------------------------------------------
        ctx->src64 = ctx->alloc<double>((size_t)N * pad4(3 * P));
        HIP_TRY(hipMemset(ctx->src64, 0, sizeof(double) * (size_t)N * pad4(3 * P)));
        ctx->ssraw = ctx->alloc<double>((size_t)N * P);
        ctx->red_len = std::max<size_t>((size_t)D * nblocks(N, 256) * 2 + 2 * (size_t)D,
                                        (size_t)kMaxBasis * kDotBlocks + kMaxBasis) + 64;
        ctx->red = ctx->alloc<double>(ctx->red_len);
        ctx->small = ctx->alloc<double>(4096);
        ctx->counter = ctx->alloc<int32_t>(16);
        ctx->spill_row = ctx->alloc<double>(16);
        HIP_TRY(hipDeviceSynchronize());
        phase("scratch");
    } catch (const std::string& e) {
        delete ctx;
        return fail(e);
    } catch (const std::exception& e) {
        delete ctx;
        return fail(e.what());
    }
    return ctx;
}

extern "C" int pbte_timing(pbte_ctx* ctx, int32_t capacity)
{
    return guarded(ctx, [&]() {
        if (capacity < 0 || capacity > (1 << 20)) throw std::string("bad timing capacity");
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        for (hipEvent_t e : ctx->events) HIP_TRY(hipEventDestroy(e));
        ctx->events.clear();
        ctx->ev_used = 0;
        for (int k = 0; k < 2 * capacity; ++k) {
            hipEvent_t e;
            HIP_TRY(hipEventCreate(&e));
            ctx->events.push_back(e);
        }
        return 0;
    });
}

extern "C" int pbte_timing_read(pbte_ctx* ctx, float* ms, int32_t n)
{
    return guarded(ctx, [&]() {
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        const int m = std::min(n, ctx->ev_used);
        for (int k = 0; k < m; ++k)
            HIP_TRY(hipEventElapsedTime(ms + k, ctx->events[2 * k], ctx->events[2 * k + 1]));
        return m;
    });
}

extern "C" int pbte_schedule_info(pbte_ctx* ctx, int64_t* out6)
{
    return guarded(ctx, [&]() {
        out6[0] = ctx->total_levels;
        out6[1] = ctx->nb;
        out6[2] = ctx->max_width;
        out6[3] = ctx->pool_peak;
        out6[4] = ctx->global_bundles;
        out6[5] = ctx->pool_slots;
        return 0;
    });
}

namespace {
template <int P>
void launch_factor(pbte_ctx* ctx, double* lu64, float* lu32)
{
    FactorArgs a{};
    a.slot_cell = ctx->slot_cell; a.bundle_ndir = ctx->bundle_ndir; a.bundle_dir_off = ctx->bundle_dir_off;
    a.bundle_dirs = ctx->bundle_dirs; a.lu_off = ctx->lu_off; a.item_off = ctx->item_off;
    a.cxv = ctx->cxv; a.cyv = ctx->cyv; a.mass = ctx->mass; a.gx = ctx->gx; a.gy = ctx->gy; a.fcm = ctx->fcm;
    a.geo = ctx->geo64; a.inv_tau_c = 1.0 / ctx->tau_c; a.N = ctx->N;
    a.lu64 = lu64; a.lu32 = lu32; a.singular = ctx->counter;
    a.refop = ctx->refop; a.cellop = ctx->cellop;
    if (ctx->refop && ctx->cellop) {
        // the padded factor storage of a bundle holds its word planes whole
        static_assert(pad4(P * P) % 4 == 0, "16-byte word planes");
        const int64_t items = (int64_t)ctx->N * ctx->max_bundle_dirs;
        dim3 grid((unsigned)((items + 63) / 64), ctx->nb);
        hipLaunchKernelGGL(factor_ref_kernel<P>, grid, dim3(64), 0, ctx->stream, a);
    } else {
        dim3 grid(nblocks(ctx->N, 128), ctx->nb);
        hipLaunchKernelGGL(factor_kernel<P>, grid, dim3(128), 0, ctx->stream, a);
    }
    HIP_TRY(hipGetLastError());
}

// the per-precision sweep data of a context
template <typename T> struct SweepData;
template <> struct SweepData<double> {
    static const double* lu(const pbte_ctx* c) { return c->lu64; }
    static const double* cx(const pbte_ctx* c) { return c->cxv; }
    static const double* cy(const pbte_ctx* c) { return c->cyv; }
    static double* crec(pbte_ctx* c) { return c->crec64; }
    static int rs(const pbte_ctx* c) { return c->rs64; }
    static double* wtr(pbte_ctx* c) { return c->wtr64; }
    static const double* kref(const pbte_ctx* c) { return c->kref64; }
};
template <> struct SweepData<float> {
    static const float* lu(const pbte_ctx* c) { return c->lu32; }
    static const float* cx(const pbte_ctx* c) { return c->cxv32; }
    static const float* cy(const pbte_ctx* c) { return c->cyv32; }
    static float* crec(pbte_ctx* c) { return c->crec32; }
    static int rs(const pbte_ctx* c) { return c->rs32; }
    static float* wtr(pbte_ctx* c) { return c->wtr32; }
    static const float* kref(const pbte_ctx* c) { return c->kref32; }
};

// LDS values of the stage slots (one face trace per slot and direction)
template <int P> size_t stage_values(const pbte_ctx* ctx, int stage)
{
    return pad4(stage * ctx->max_bundle_dirs * Tr<P>::TQ);
}

// LDS bytes of the pipelined kernel: pool, zero and spill slots, stage
// slots, wall traces
template <int P, typename T> size_t sweep_lds_bytes(const pbte_ctx* ctx, int stage = 0)
{
    return ((size_t)(ctx->pool_slots + 2) * ctx->max_bundle_dirs * Tr<P>::W + stage_values<P>(ctx, stage) +
            (size_t)ctx->max_wall_in * ctx->max_bundle_dirs * pad4(P)) *
           sizeof(T);
}

// stage slots a pipelined launch holds (all the context's staged reads, or
// none when they do not fit beside the rest of its LDS, extra bytes
// included: its staged faces then read global memory)
template <int P, typename T> int stage_slots(const pbte_ctx* ctx, int use_pbuf, size_t extra = 0)
{
    const int n = ctx->stage_slots[use_pbuf ? 1 : 0];
    return ctx->staging && n > 0 && sweep_lds_bytes<P, T>(ctx, n) + extra <= ctx->lds_blk ? n : 0;
}

// whether a pipelined launch reads global memory for upwind values (the
// kernel instance with the global-read branch)
bool needs_gread(const pbte_ctx* ctx, int use_pbuf, bool stage_on)
{
    if (ctx->global_bundles > 0) return true;
    const int64_t np_old = stage_on ? ctx->np_old_unstaged : ctx->np_old_reads;
    const int64_t per = use_pbuf ? (stage_on ? ctx->per_unstaged : ctx->periodic_reads)
                                 : (stage_on ? ctx->per_old_unstaged : ctx->per_old_reads);
    return np_old + per > 0;
}

template <int P, typename T, int PF, bool GR, int DEPTH>
void launch_pipe(pbte_ctx* ctx, const SweepArgs<T>& a, size_t lds)
{
    auto* fn = ctx->max_bundle_dirs == 1 ? sweep_pipe_kernel<P, T, PF, GR, DEPTH, 1>
                                         : sweep_pipe_kernel<P, T, PF, GR, DEPTH, 0>;
    if (lds > 65536)
        HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(fn), hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds));
    hipLaunchKernelGGL(fn, dim3(ctx->nb), dim3(ctx->block), lds, ctx->stream, a);
}

// the kernel arguments of one context's sweep
template <int P, typename T>
SweepArgs<T> sweep_args(pbte_ctx* ctx, const T* pbuf, T* vdf, int use_pbuf, int homogeneous, int stage = 0)
{
    using S = SweepData<T>;
    SweepArgs<T> a{};
    a.block_bundle = ctx->block_bundle; a.bundle_ndir = ctx->bundle_ndir; a.bundle_dir_off = ctx->bundle_dir_off;
    a.bundle_dirs = ctx->bundle_dirs; a.lev_off = ctx->lev_off; a.lev_ptr = ctx->lev_ptr;
    a.lu_off = ctx->lu_off; a.item_off = ctx->item_off;
    a.lu = S::lu(ctx); a.cxv = S::cx(ctx); a.cyv = S::cy(ctx);
    a.crec = S::crec(ctx); a.rs = S::rs(ctx); a.other = ctx->other; a.wtr = S::wtr(ctx);
    a.wcell = ctx->wcell; a.wc_off = ctx->wc_off;
    a.pbuf = pbuf; a.vdf = vdf; a.N = ctx->N; a.F = ctx->F; a.use_pbuf = use_pbuf; a.homogeneous = homogeneous;
    a.irec = ctx->irec;
    a.spill_row = reinterpret_cast<T*>(ctx->spill_row);
    a.bundle_global = ctx->bundle_global;
    a.stage_list = ctx->stage_list; a.stage_off = ctx->stage_off; a.stage_on = stage > 0;
    a.lane_zero = 0;
    a.pool = ctx->pool_slots;
    a.wbase = (ctx->pool_slots + 2) * ctx->max_bundle_dirs * Tr<P>::W + (int)stage_values<P>(ctx, stage);
    a.fbase = a.wbase + ctx->max_wall_in * ctx->max_bundle_dirs * pad4(P);
    a.nthr = ctx->block;
    for (int k = 0; k < ctx->Q * ctx->Q; ++k) a.kref[k] = S::kref(ctx)[k];
    a.mass = ctx->mass; a.gx = ctx->gx; a.gy = ctx->gy; a.fcm = ctx->fcm; a.geo = ctx->geo64;
    a.cxv64 = ctx->cxv; a.cyv64 = ctx->cyv; a.inv_tau_c = 1.0 / ctx->tau_c;
    a.refop = ctx->refop; a.cellop = ctx->cellop; a.lu32 = ctx->lu32;
    return a;
}

// whether a context's sweep of this precision runs the pipelined kernel
// the pipelined kernel addresses a bundle's factors through a buffer
// descriptor (32-bit byte range and offsets); a bundle whose factors exceed
// that runs the plain level kernel, which indexes with 64-bit pointers
template <typename T> bool factors_fit_descriptor(const pbte_ctx* ctx)
{
    for (size_t k = 0; k + 1 < ctx->h_lu_off.size(); ++k)
        if ((uint64_t)(ctx->h_lu_off[k + 1] - ctx->h_lu_off[k]) * sizeof(T) > 0xffffffffull) return false;
    return true;
}

template <int P, typename T> bool pipelined_shape(const pbte_ctx* ctx)
{
    return ctx->variant != 1 && ctx->block <= PipeThreads<T>::max && sweep_lds_bytes<P, T>(ctx) <= ctx->lds_blk;
}

template <int P, typename T> bool runs_pipelined(const pbte_ctx* ctx)
{
    return pipelined_shape<P, T>(ctx) && factors_fit_descriptor<T>(ctx);
}

// a binary32 sweep of this context runs the fused kernel
template <int P, typename T> bool runs_fused(const pbte_ctx* ctx)
{
    return std::is_same<T, float>::value && P <= 10 && ctx->fused32 && pipelined_shape<P, T>(ctx);
}

// LDS bytes of the split fused sweep: the pipelined kernel's plus two
// factor buffers of the consumer width
template <int P, typename T> size_t split_lds_bytes(const pbte_ctx* ctx)
{
    return sweep_lds_bytes<P, T>(ctx) + (size_t)2 * fac_vals<P>() * ctx->block * sizeof(T);
}

// ... and runs it as the split kernel (producer and consumer halves)
template <int P, typename T> bool runs_split(const pbte_ctx* ctx)
{
    return P <= 6 && runs_fused<P, T>(ctx) && ctx->split && ctx->block % 64 == 0 && 2 * ctx->block <= 512 &&
           split_lds_bytes<P, T>(ctx) <= ctx->lds_blk;
}

template <int P, typename T, bool GR>
void launch_split(pbte_ctx* ctx, const SweepArgs<T>& a, size_t lds)
{
    auto* fn = ctx->max_bundle_dirs == 1 ? sweep_split_kernel<P, T, GR, 1> : sweep_split_kernel<P, T, GR, 0>;
    if (lds > 65536)
        HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(fn), hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds));
    hipLaunchKernelGGL(fn, dim3(ctx->nb), dim3(2 * ctx->block), lds, ctx->stream, a);
}

template <int P, typename T>
void launch_sweep(pbte_ctx* ctx, const T* pbuf, T* vdf, int use_pbuf, int homogeneous, bool onthefly)
{
    // the split kernel's factor buffers follow the rest of its LDS (a
    // fused context runs the fused kernels on either entry point)
    bool split = false;
    if constexpr (P <= 6) split = runs_split<P, T>(ctx);
    const size_t extra = split ? split_lds_bytes<P, T>(ctx) - sweep_lds_bytes<P, T>(ctx) : 0;
    const bool piped = runs_fused<P, T>(ctx) || (!onthefly && runs_pipelined<P, T>(ctx));
    const int stage = piped ? stage_slots<P, T>(ctx, use_pbuf, extra) : 0;
    const SweepArgs<T> a = sweep_args<P, T>(ctx, pbuf, vdf, use_pbuf, homogeneous, stage);
    const size_t lds = sweep_lds_bytes<P, T>(ctx, stage);
    const bool gread = needs_gread(ctx, use_pbuf, stage > 0);
    const bool timed = 2 * (size_t)ctx->ev_used + 1 < ctx->events.size();
    if (timed) HIP_TRY(hipEventRecord(ctx->events[2 * ctx->ev_used], ctx->stream));
    if (runs_fused<P, T>(ctx)) {
        // binary32 with the element operators built and factorised in the
        // sweep (stored binary32 factors, if any, are not read)
        if constexpr (std::is_same<T, float>::value && P <= 10) {
            if constexpr (P <= 6) {
                if (split) {
                    if (gread) launch_split<P, T, true>(ctx, a, lds + extra);
                    else launch_split<P, T, false>(ctx, a, lds + extra);
                }
            }
            if (!split && ctx->depth == 4) {
                if (gread) launch_pipe<P, T, LU_FUSED, true, 4>(ctx, a, lds);
                else launch_pipe<P, T, LU_FUSED, false, 4>(ctx, a, lds);
            } else if (!split && ctx->depth == 3) {
                if (gread) launch_pipe<P, T, LU_FUSED, true, 3>(ctx, a, lds);
                else launch_pipe<P, T, LU_FUSED, false, 3>(ctx, a, lds);
            } else if (!split) {
                if (gread) launch_pipe<P, T, LU_FUSED, true, 2>(ctx, a, lds);
                else launch_pipe<P, T, LU_FUSED, false, 2>(ctx, a, lds);
            }
        }
    } else if (onthefly) {
        bool refined = false;
        if constexpr (std::is_same<T, double>::value && P <= 10) {
            // binary32 inverses stored (the "mixed" mode): refine from them
            // instead of building and factorising every operator
            if (ctx->have32 && ctx->refop) {
                hipLaunchKernelGGL((sweep_refine_kernel<P>), dim3(ctx->nb), dim3(std::min(ctx->block, 256)),
                                   6 * P * P * sizeof(double), ctx->stream, a);
                refined = true;
            }
        }
        if (!refined)
            hipLaunchKernelGGL((sweep_otf_kernel<P, T>), dim3(ctx->nb), dim3(std::min(ctx->block, 256)), 0, ctx->stream, a);
    } else if (!runs_pipelined<P, T>(ctx)) {
        // the plain level kernel: asked for, or a binary64 sweep whose pool,
        // sized for binary32, exceeds a workgroup's LDS
        hipLaunchKernelGGL((sweep_kernel<P, T>), dim3(ctx->nb), dim3(std::min(ctx->block, 256)), 0, ctx->stream, a);
    } else {
        constexpr int pf = lu_stored_mode<P, T>();
        // (the four-level pipeline is built for the fused sweep only)
        if (ctx->depth >= 3) {
            if (gread) launch_pipe<P, T, pf, true, 3>(ctx, a, lds);
            else launch_pipe<P, T, pf, false, 3>(ctx, a, lds);
        } else {
            if (gread) launch_pipe<P, T, pf, true, 2>(ctx, a, lds);
            else launch_pipe<P, T, pf, false, 2>(ctx, a, lds);
        }
    }
    HIP_TRY(hipGetLastError());
    if (timed) {
        HIP_TRY(hipEventRecord(ctx->events[2 * ctx->ev_used + 1], ctx->stream));
        ++ctx->ev_used;
    }
}

template <typename T>
void launch_source(pbte_ctx* ctx, const double* ts, const double* qxs, const double* qys, T* out, int stride);

// one sweep: the source into the cell records, the adiabatic wall vectors
// from flux_wall, then the sweep
template <typename T>
void prep_sweep(pbte_ctx* ctx, const double* ts, const double* qxs, const double* qys, const double* flux_wall)
{
    using S = SweepData<T>;
    const int P = ctx->P;
    launch_source<T>(ctx, ts, qxs, qys, S::crec(ctx), S::rs(ctx));
    if (ctx->n_adw) {
        const unsigned g = nblocks((int64_t)ctx->n_adw * P, 256);
#define WT(PP) case PP: hipLaunchKernelGGL((wall_vec_kernel<PP, T>), dim3(g), dim3(256), 0, ctx->stream, \
                                           flux_wall, ctx->adw_row, ctx->adw_face, ctx->adw_rlen, ctx->n_adw, \
                                           S::wtr(ctx)); break;
        switch (P) { WT(3) WT(6) WT(10) WT(15) }
#undef WT
        HIP_TRY(hipGetLastError());
    }
}

template <typename T>
void run_sweep(pbte_ctx* ctx, const double* ts, const double* qxs, const double* qys, const double* flux_wall,
               const T* pbuf, T* vdf, int use_pbuf, int homogeneous, bool onthefly = false)
{
    const int P = ctx->P;
    prep_sweep<T>(ctx, ts, qxs, qys, flux_wall);
    switch (P) {
#define SW(PP) case PP: launch_sweep<PP, T>(ctx, pbuf, vdf, use_pbuf, homogeneous, onthefly); break;
    SW(3) SW(6) SW(10) SW(15)
#undef SW
    }
}

template <typename T>
void launch_source(pbte_ctx* ctx, const double* ts, const double* qxs, const double* qys, T* out, int stride)
{
    const int P = ctx->P;
    const unsigned g = nblocks((int64_t)ctx->N * P, 256);
#define SRC_CASE(PP)                                                                                 \
    case PP:                                                                                         \
        hipLaunchKernelGGL((source_kernel<PP, T>), dim3(g), dim3(256), 0, ctx->stream, ctx->mass, ts, qxs, \
                           qys, out, stride, ctx->N, ctx->cv, ctx->pi, ctx->tau_r, ctx->tau_n, ctx->vg); \
        break;
    switch (P) { SRC_CASE(3) SRC_CASE(6) SRC_CASE(10) SRC_CASE(15) }
#undef SRC_CASE
    HIP_TRY(hipGetLastError());
}
}  // namespace

extern "C" int pbte_factorise(pbte_ctx* ctx, int32_t precision)
{
    return guarded(ctx, [&]() {
        if (precision != PBTE_FP64 && precision != PBTE_FP32 && precision != PBTE_FP64_FP32)
            throw std::string("bad precision");
        double* l64 = nullptr;
        float* l32 = nullptr;
        if (precision != PBTE_FP32) {
            if (!ctx->lu64) ctx->lu64 = ctx->alloc<double>((size_t)ctx->lu_elems);
            l64 = ctx->lu64;
        }
        if (precision != PBTE_FP64) {
            if (!ctx->lu32) ctx->lu32 = ctx->alloc<float>((size_t)ctx->lu_elems);
            l32 = ctx->lu32;
        }
        HIP_TRY(hipMemsetAsync(ctx->counter, 0, sizeof(int32_t), ctx->stream));
        switch (ctx->P) {
        case 3: launch_factor<3>(ctx, l64, l32); break;
        case 6: launch_factor<6>(ctx, l64, l32); break;
        case 10: launch_factor<10>(ctx, l64, l32); break;
        case 15: launch_factor<15>(ctx, l64, l32); break;
        }
        int32_t bad = 0;
        HIP_TRY(hipMemcpyAsync(&bad, ctx->counter, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        if (l64) ctx->have64 = true;
        if (l32) ctx->have32 = true;
        return (int)bad;
    });
}

extern "C" int pbte_sweep(pbte_ctx* ctx, int32_t precision, const double* ts, const double* qxs,
                          const double* qys, void* vdf, const double* flux_wall, const void* pbuf,
                          int32_t use_pbuf, int32_t homogeneous)
{
    return guarded(ctx, [&]() {
        if (use_pbuf && !pbuf) throw std::string("use_pbuf without a buffer");
        if (ctx->n_wall && !flux_wall) throw std::string("adiabatic walls need flux_wall");
        if (precision == PBTE_FP64) {
            if (!ctx->have64) throw std::string("FP64 factors not built; call pbte_factorise");
            run_sweep<double>(ctx, ts, qxs, qys, flux_wall, (const double*)pbuf, (double*)vdf, use_pbuf,
                              homogeneous);
        } else if (precision == PBTE_FP32) {
            if (!ctx->have32 && !ctx->fused32)
                throw std::string("FP32 factors not built; call pbte_factorise or pbte_set_fused");
            run_sweep<float>(ctx, ts, qxs, qys, flux_wall, (const float*)pbuf, (float*)vdf, use_pbuf,
                             homogeneous);
        } else {
            throw std::string("bad precision");
        }
        return 0;
    });
}

namespace {
// the context's reference-element matrices and unit-edge matrix against
// the fused sweep's compile-time tables
template <int P> bool fused_tables_match(const std::vector<double>& f)
{
    using R = RefElem<P>;
    constexpr int PP = P * P, Q = Deg<P>::v + 1;
    double big = 0.0, err = 0.0;
    for (int k = 0; k < PP; ++k) {
        const double t[3] = {R::M[k], R::Dx[k], R::Dy[k]};
        for (int j = 0; j < 3; ++j) {
            big = std::max(big, std::fabs(t[j]));
            err = std::max(err, std::fabs(f[(size_t)j * PP + k] - t[j]));
        }
    }
    for (int r = 0; r < Q; ++r)
        for (int q = 0; q < Q; ++q)
            err = std::max(err, std::fabs(f[(size_t)3 * PP + r * Q + q] - unit_edge<P>(r, q)));
    return big > 0.0 && err <= 1e-12 * big;
}
}  // namespace

extern "C" int pbte_set_cell_operators(pbte_ctx* ctx, const double* ref, const double* cell)
{
    return guarded(ctx, [&]() {
        if (!ref || !cell) throw std::string("null operator arrays");
        const size_t nr = (size_t)6 * ctx->P * ctx->P, nc = (size_t)ctx->N * 8;
        if (!ctx->refop) ctx->refop = ctx->alloc<double>(nr);
        if (!ctx->cellop) ctx->cellop = ctx->alloc<double>(nc);
        HIP_TRY(hipMemcpy(ctx->refop, ref, nr * sizeof(double), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(ctx->cellop, cell, nc * sizeof(double), hipMemcpyHostToDevice));
        // the fused sweep's data: the reference-element mass and derivative
        // matrices, and per element area2 / tau_c and area2 times the
        // gradient coefficients (the element tensors are these products:
        // mass = area2 M, gx = area2 (a D_xi + b D_eta), gy = area2 (c D_xi
        // + d D_eta)), written into the padding of the cell records
        const int P = ctx->P, PP = P * P;
        for (int k = 0; k < 3 * PP; ++k) ctx->h_fref[k] = ref[k];
        bool tables = false;
        switch (P) {
        case 3: tables = fused_tables_match<3>(ctx->h_fref); break;
        case 6: tables = fused_tables_match<6>(ctx->h_fref); break;
        case 10: tables = fused_tables_match<10>(ctx->h_fref); break;
        }
        ctx->fused_tables = tables;
        HIP_TRY(hipMemsetAsync(ctx->counter, 0, sizeof(int32_t), ctx->stream));
        hipLaunchKernelGGL(fused_scalars_kernel, dim3(nblocks(ctx->N, 256)), dim3(256), 0, ctx->stream, ctx->cellop,
                           ctx->tau_c, ctx->N, 3 * P + 9, 3 * P, ctx->crec64, ctx->rs64, ctx->crec32, ctx->rs32,
                           ctx->counter);
        HIP_TRY(hipGetLastError());
        int32_t bad = 0;
        HIP_TRY(hipMemcpyAsync(&bad, ctx->counter, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        ctx->fused_geometry = bad == 0;
        ctx->have_cellops = true;
        return 0;
    });
}

extern "C" int pbte_sweep_path(pbte_ctx* ctx, int32_t precision)
{
    return guarded(ctx, [&]() {
        if (precision != PBTE_FP64 && precision != PBTE_FP32) throw std::string("bad precision");
        const bool f32 = precision == PBTE_FP32;
        const bool stored = f32 ? ctx->have32 : ctx->have64;
        int path = 0;
        switch (ctx->P) {
#define SP(PP)                                                                                                  \
    case PP:                                                                                                    \
        if (f32 && runs_split<PP, float>(ctx)) path = 4;                                                        \
        else if (f32 && runs_fused<PP, float>(ctx)) path = 3;                                                   \
        else if (!stored) path = f32 ? 5 : (ctx->have32 && ctx->refop && PP <= 10 ? 6 : 5);                     \
        else path = (f32 ? runs_pipelined<PP, float>(ctx) : runs_pipelined<PP, double>(ctx)) ? 2 : 1;           \
        break;
            SP(3) SP(6) SP(10) SP(15)
#undef SP
        }
        return path;
    });
}

extern "C" int pbte_staging(pbte_ctx* ctx, int32_t on, int64_t* out4)
{
    return guarded(ctx, [&]() {
        if (on > 1) throw std::string("staging: on must be -1, 0 or 1");
        if (on >= 0) ctx->staging = on != 0;
        if (out4) {
            out4[0] = ctx->stage_slots[0];
            out4[1] = ctx->stage_slots[1];
            out4[2] = ctx->np_old_reads + ctx->per_old_reads - ctx->np_old_unstaged - ctx->per_old_unstaged;
            out4[3] = ctx->periodic_reads - ctx->per_unstaged;
        }
        return ctx->staging ? 1 : 0;
    });
}

extern "C" int pbte_set_fused(pbte_ctx* ctx, int32_t on)
{
    return guarded(ctx, [&]() {
        if (!on) {
            ctx->fused32 = false;
            return 0;
        }
        if (on != 1 && on != 2) throw std::string("fused sweep: on is 0, 1 or 2");
        ctx->split = on == 2;
        if (ctx->P > 10) throw std::string("fused sweep: element degree above 3");
        if (!ctx->have_cellops) throw std::string("fused sweep: pbte_set_cell_operators first");
        if (!ctx->fused_ok)
            throw std::string("fused sweep: face mass blocks are not a length times one unit-edge matrix");
        if (!ctx->fused_tables)
            throw std::string("fused sweep: the reference-element matrices differ from the compiled tables");
        if (!ctx->fused_geometry)
            throw std::string("fused sweep: the gradient coefficients are not the faces' normals (face order?)");
        bool shape = false;
        switch (ctx->P) {
        case 3: shape = pipelined_shape<3, float>(ctx); break;
        case 6: shape = pipelined_shape<6, float>(ctx); break;
        case 10: shape = pipelined_shape<10, float>(ctx); break;
        }
        if (!shape) throw std::string("fused sweep: the context does not run the pipelined kernel");
        ctx->fused32 = true;
        return 0;
    });
}

extern "C" int pbte_sweep_onthefly(pbte_ctx* ctx, int32_t precision, const double* ts, const double* qxs,
                                   const double* qys, void* vdf, const double* flux_wall, const void* pbuf,
                                   int32_t use_pbuf, int32_t homogeneous)
{
    return guarded(ctx, [&]() {
        if (use_pbuf && !pbuf) throw std::string("use_pbuf without a buffer");
        if (ctx->n_wall && !flux_wall) throw std::string("adiabatic walls need flux_wall");
        if (precision == PBTE_FP64)
            run_sweep<double>(ctx, ts, qxs, qys, flux_wall, (const double*)pbuf, (double*)vdf, use_pbuf,
                              homogeneous, true);
        else if (precision == PBTE_FP32)
            run_sweep<float>(ctx, ts, qxs, qys, flux_wall, (const float*)pbuf, (float*)vdf, use_pbuf,
                             homogeneous, true);
        else
            throw std::string("bad precision");
        return 0;
    });
}

// ==========================================================================
// batched sweeps: the bundles of several contexts (independent problems of
// one element degree, built on one stream) in one persistent launch of the
// pipelined kernel, whose workgroups take (context, bundle) work items from
// a device queue, longest first
// ==========================================================================
struct pbte_batch {
    std::vector<pbte_ctx*> ctxs;
    int2* bmap = nullptr;           // device [blocks] {context, block}: the work items, longest first
    void* args = nullptr;           // device SweepArgs<T>[n] of the next launch (an entry of the cache)
    int* head = nullptr;            // device work-queue head
    // argument sets already on the device (a lockstep iteration alternates
    // between a few distribution buffers): never rewritten, so a launch
    // still reading one is never disturbed; least recently used dropped
    struct ArgSet { std::vector<unsigned char> host; void* dev; uint64_t used; };
    std::vector<ArgSet> arg_cache;
    uint64_t arg_clock = 0;
    static constexpr size_t kArgSets = 8;
    // the batch's own stream: its launch waits for the contexts' stream up
    // to the sweep's preparation, and the contexts' stream waits for the
    // launch at the join, so the launches of several batches overlap
    hipStream_t stream = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    bool pending = false;           // a launch the contexts' stream has not joined yet
    int nblk = 0, block = 0, P = 0, ncu = 0;
    std::string err;
    ~pbte_batch()
    {
        if (pending && !ctxs.empty()) (void)hipStreamWaitEvent(ctxs[0]->stream, ev_out, 0);
        if (stream) (void)hipStreamSynchronize(stream);
        if (bmap) (void)hipFree(bmap);
        for (auto& e : arg_cache) (void)hipFree(e.dev);
        if (head) (void)hipFree(head);
        if (ev_in) (void)hipEventDestroy(ev_in);
        if (ev_out) (void)hipEventDestroy(ev_out);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

extern "C" pbte_batch* pbte_batch_create(pbte_ctx* const* ctxs, int32_t n, char* err, int32_t errlen)
{
    auto fail = [&](const std::string& m) -> pbte_batch* {
        if (err && errlen > 0) {
            std::strncpy(err, m.c_str(), (size_t)errlen - 1);
            err[errlen - 1] = 0;
        }
        return nullptr;
    };
    if (!ctxs || n <= 0) return fail("empty batch");
    auto* bt = new pbte_batch();
    try {
        std::vector<int2> h;
        std::vector<int> by_len(n);
        for (int k = 0; k < n; ++k) {
            pbte_ctx* c = ctxs[k];
            if (!c) throw std::string("null context in batch");
            if (c->P != ctxs[0]->P) throw std::string("a batch needs one element degree");
            if (c->stream != ctxs[0]->stream) throw std::string("a batch needs one stream");
            if (c->depth != ctxs[0]->depth) throw std::string("a batch needs one pipeline depth");
            bt->ctxs.push_back(c);
            bt->block = std::max(bt->block, c->block);
            by_len[k] = k;
        }
        // longest bundles first (levels per bundle, the workgroup's run
        // time): workgroups start in block order, so the short ones fill the
        // CUs the long ones leave; a context's blocks stay together, in its
        // own XCD-dealt order
        auto len = [&](int k) { return (double)ctxs[k]->total_levels / std::max(ctxs[k]->nb, 1); };
        std::stable_sort(by_len.begin(), by_len.end(), [&](int a, int b) { return len(a) > len(b); });
        for (int k : by_len)
            for (int b = 0; b < ctxs[k]->nb; ++b) h.push_back(make_int2(k, b));
        // contexts with narrower workgroups run in the widest one: their
        // extra threads are idle items of every level
        bt->P = ctxs[0]->P;
        bt->nblk = (int)h.size();
        HIP_TRY(hipMalloc((void**)&bt->bmap, h.size() * sizeof(int2)));
        HIP_TRY(hipMemcpy(bt->bmap, h.data(), h.size() * sizeof(int2), hipMemcpyHostToDevice));
        HIP_TRY(hipMalloc((void**)&bt->head, sizeof(int)));
        HIP_TRY(hipStreamCreateWithFlags(&bt->stream, hipStreamNonBlocking));
        HIP_TRY(hipEventCreateWithFlags(&bt->ev_in, hipEventDisableTiming));
        HIP_TRY(hipEventCreateWithFlags(&bt->ev_out, hipEventDisableTiming));
        int dev = 0;
        HIP_TRY(hipGetDevice(&dev));
        HIP_TRY(hipDeviceGetAttribute(&bt->ncu, hipDeviceAttributeMultiprocessorCount, dev));
        return bt;
    } catch (const std::string& m) {
        delete bt;
        return fail(m);
    }
}

extern "C" void pbte_batch_destroy(pbte_batch* bt) { delete bt; }

extern "C" int pbte_batch_pipelined(pbte_batch* bt, int32_t precision)
{
    if (!bt) return -1;
    for (pbte_ctx* c : bt->ctxs) {
        bool ok = false;
        switch (bt->P) {
#define BP(PP) case PP: ok = precision == PBTE_FP64 ? (c->have64 && runs_pipelined<PP, double>(c)) \
                                                    : (runs_fused<PP, float>(c) == runs_fused<PP, float>(bt->ctxs[0]) && \
                                                       (runs_fused<PP, float>(c) || (c->have32 && runs_pipelined<PP, float>(c)))); \
                        break;
            BP(3) BP(6) BP(10) BP(15)
#undef BP
        }
        if (!ok) return 0;
    }
    return 1;
}

extern "C" const char* pbte_batch_last_error(pbte_batch* bt) { return bt ? bt->err.c_str() : "null batch"; }

template <int P, typename T, int PF, bool GR, int DEPTH>
void launch_pipe_batch(pbte_batch* bt, bool one_dir, size_t lds, hipStream_t st)
{
    auto* fn = one_dir ? sweep_queue_kernel<P, T, PF, GR, DEPTH, 1> : sweep_queue_kernel<P, T, PF, GR, DEPTH, 0>;
    if (lds > 65536)
        HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(fn), hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds));
    // as many workgroups as the CUs hold at once (registers and LDS), at
    // most one per work item
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(fn), bt->block, lds) !=
        hipSuccess) {
        (void)hipGetLastError();
        per_cu = 1;
    }
    const int grid = std::max(1, std::min(bt->nblk, std::max(per_cu, 1) * std::max(bt->ncu, 1)));
    hipLaunchKernelGGL(queue_reset_kernel, dim3(1), dim3(1), 0, st, bt->head);
    hipLaunchKernelGGL(fn, dim3(grid), dim3(bt->block), lds, st, static_cast<const SweepArgs<T>*>(bt->args),
                       bt->bmap, bt->nblk, bt->head);
}

template <int P, typename T>
void batch_sweep(pbte_batch* bt, T* const* vdf, const void* const* pbuf, int homogeneous)
{
    const int n = (int)bt->ctxs.size();
    std::vector<unsigned char> host((size_t)n * sizeof(SweepArgs<T>));
    size_t lds = 0;
    bool gread = false, one_dir = true;
    const bool fused = runs_fused<P, T>(bt->ctxs[0]);
    for (int k = 0; k < n; ++k) {
        pbte_ctx* c = bt->ctxs[k];
        if (runs_fused<P, T>(c) != fused)
            throw std::string("batch: fused and stored-factor binary32 contexts in one launch");
        if (!fused && !runs_pipelined<P, T>(c))
            throw std::string("batch: a context does not run the pipelined kernel in this precision");
        const T* pb = pbuf ? static_cast<const T*>(pbuf[k]) : nullptr;
        const int use_pbuf = pb != nullptr;
        const int stage = stage_slots<P, T>(c, use_pbuf);
        const SweepArgs<T> a = sweep_args<P, T>(c, pb, vdf[k], use_pbuf, homogeneous, stage);
        std::memcpy(host.data() + (size_t)k * sizeof(SweepArgs<T>), &a, sizeof(a));
        lds = std::max(lds, sweep_lds_bytes<P, T>(c, stage));
        gread = gread || needs_gread(c, use_pbuf, stage > 0);
        one_dir = one_dir && c->max_bundle_dirs == 1;
    }
    // the arguments change only with the distributions and the source flag:
    // a set seen before is on the device already
    bt->args = nullptr;
    for (auto& e : bt->arg_cache)
        if (e.host == host) {
            bt->args = e.dev;
            e.used = ++bt->arg_clock;
        }
    if (!bt->args) {
        if (bt->arg_cache.size() >= pbte_batch::kArgSets) {
            // the least recently used set: its last launch has finished once
            // the batch stream is idle
            auto old = std::min_element(bt->arg_cache.begin(), bt->arg_cache.end(),
                                        [](const auto& a, const auto& b) { return a.used < b.used; });
            HIP_TRY(hipStreamSynchronize(bt->stream));
            HIP_TRY(hipFree(old->dev));
            bt->arg_cache.erase(old);
        }
        void* dev = nullptr;
        HIP_TRY(hipMalloc(&dev, host.size()));
        HIP_TRY(hipMemcpy(dev, host.data(), host.size(), hipMemcpyHostToDevice));
        bt->arg_cache.push_back({host, dev, ++bt->arg_clock});
        bt->args = dev;
    }
    constexpr int pf = lu_stored_mode<P, T>();
    // the launch follows the contexts' stream up to here (sources, wall
    // vectors) on the batch's stream
    HIP_TRY(hipEventRecord(bt->ev_in, bt->ctxs[0]->stream));
    HIP_TRY(hipStreamWaitEvent(bt->stream, bt->ev_in, 0));
    const hipStream_t st = bt->stream;
    if (fused) {
        if constexpr (std::is_same<T, float>::value && P <= 10) {
            if (bt->ctxs[0]->depth == 4) {
                if (gread) launch_pipe_batch<P, T, LU_FUSED, true, 4>(bt, one_dir, lds, st);
                else launch_pipe_batch<P, T, LU_FUSED, false, 4>(bt, one_dir, lds, st);
            } else if (bt->ctxs[0]->depth == 3) {
                if (gread) launch_pipe_batch<P, T, LU_FUSED, true, 3>(bt, one_dir, lds, st);
                else launch_pipe_batch<P, T, LU_FUSED, false, 3>(bt, one_dir, lds, st);
            } else {
                if (gread) launch_pipe_batch<P, T, LU_FUSED, true, 2>(bt, one_dir, lds, st);
                else launch_pipe_batch<P, T, LU_FUSED, false, 2>(bt, one_dir, lds, st);
            }
        }
    } else if (bt->ctxs[0]->depth >= 3) {
        if (gread) launch_pipe_batch<P, T, pf, true, 3>(bt, one_dir, lds, st);
        else launch_pipe_batch<P, T, pf, false, 3>(bt, one_dir, lds, st);
    } else {
        if (gread) launch_pipe_batch<P, T, pf, true, 2>(bt, one_dir, lds, st);
        else launch_pipe_batch<P, T, pf, false, 2>(bt, one_dir, lds, st);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(bt->ev_out, bt->stream));
    bt->pending = true;
}

// the contexts' stream waits for the batch's last launch
void batch_join(pbte_batch* bt)
{
    if (!bt->pending) return;
    HIP_TRY(hipStreamWaitEvent(bt->ctxs[0]->stream, bt->ev_out, 0));
    bt->pending = false;
}

extern "C" int pbte_batch_join(pbte_batch* bt)
{
    if (!bt) return -1;
    try {
        batch_join(bt);
        return 0;
    } catch (const std::string& m) {
        bt->err = m;
        return -1;
    }
}

extern "C" int pbte_batch_sweep_async(pbte_batch* bt, int32_t precision, const double* const* ts,
                                      const double* const* qxs, const double* const* qys, void* const* vdf,
                                      const double* const* flux_wall, const void* const* pbuf, int32_t homogeneous)
{
    if (!bt) return -1;
    try {
        // a previous launch of this batch still reads its contexts' cell
        // records, which the preparation below rewrites
        batch_join(bt);
        const int n = (int)bt->ctxs.size();
        for (int k = 0; k < n; ++k) {
            pbte_ctx* c = bt->ctxs[k];
            if (c->n_wall && !flux_wall[k]) throw std::string("adiabatic walls need flux_wall");
            if (precision == PBTE_FP64) {
                if (!c->have64) throw std::string("FP64 factors not built in a batch context");
                prep_sweep<double>(c, ts[k], qxs[k], qys[k], flux_wall[k]);
            } else if (precision == PBTE_FP32) {
                if (!c->have32 && !c->fused32) throw std::string("FP32 factors not built in a batch context");
                prep_sweep<float>(c, ts[k], qxs[k], qys[k], flux_wall[k]);
            } else {
                throw std::string("bad precision");
            }
        }
        if (precision == PBTE_FP64) {
            auto v = reinterpret_cast<double* const*>(vdf);
            switch (bt->P) {
            case 3: batch_sweep<3, double>(bt, v, pbuf, homogeneous); break;
            case 6: batch_sweep<6, double>(bt, v, pbuf, homogeneous); break;
            case 10: batch_sweep<10, double>(bt, v, pbuf, homogeneous); break;
            case 15: batch_sweep<15, double>(bt, v, pbuf, homogeneous); break;
            }
        } else {
            auto v = reinterpret_cast<float* const*>(vdf);
            switch (bt->P) {
            case 3: batch_sweep<3, float>(bt, v, pbuf, homogeneous); break;
            case 6: batch_sweep<6, float>(bt, v, pbuf, homogeneous); break;
            case 10: batch_sweep<10, float>(bt, v, pbuf, homogeneous); break;
            case 15: batch_sweep<15, float>(bt, v, pbuf, homogeneous); break;
            }
        }
        return 0;
    } catch (const std::string& m) {
        bt->err = m;
        return -1;
    }
}

extern "C" int pbte_batch_sweep(pbte_batch* bt, int32_t precision, const double* const* ts, const double* const* qxs,
                                const double* const* qys, void* const* vdf, const double* const* flux_wall,
                                const void* const* pbuf, int32_t homogeneous)
{
    const int st = pbte_batch_sweep_async(bt, precision, ts, qxs, qys, vdf, flux_wall, pbuf, homogeneous);
    return st < 0 ? st : pbte_batch_join(bt);
}

extern "C" int pbte_moments(pbte_ctx* ctx, int32_t precision, const void* vdf, double* ts, double* qxs,
                            double* qys, double* temp, double* qx, double* qy)
{
    return guarded(ctx, [&]() {
        const int64_t np = (int64_t)ctx->N * ctx->P;
        const bool aligned = (reinterpret_cast<uintptr_t>(vdf) & 15u) == 0;
        const unsigned g = nblocks(np, 256);
        if (precision == PBTE_FP64) {
            if (aligned && np % 2 == 0)
                hipLaunchKernelGGL((moments_vec_kernel<double, 2>), dim3(nblocks(np / 2, 256)), dim3(256), 0,
                                   ctx->stream, (const double*)vdf, ctx->cxv, ctx->cyv, ctx->domega, ctx->D, ctx->N,
                                   ctx->P, ctx->cv, ts, qxs, qys, ctx->ssraw);
            else
                hipLaunchKernelGGL(moments_kernel<double>, dim3(g), dim3(256), 0, ctx->stream, (const double*)vdf,
                                   ctx->cxv, ctx->cyv, ctx->domega, ctx->D, ctx->N, ctx->P, ctx->cv, ts, qxs, qys,
                                   ctx->ssraw);
        } else {
            if (aligned && np % 4 == 0)
                hipLaunchKernelGGL((moments_vec_kernel<float, 4>), dim3(nblocks(np / 4, 256)), dim3(256), 0,
                                   ctx->stream, (const float*)vdf, ctx->cxv, ctx->cyv, ctx->domega, ctx->D, ctx->N,
                                   ctx->P, ctx->cv, ts, qxs, qys, ctx->ssraw);
            else
                hipLaunchKernelGGL(moments_kernel<float>, dim3(g), dim3(256), 0, ctx->stream, (const float*)vdf,
                                   ctx->cxv, ctx->cyv, ctx->domega, ctx->D, ctx->N, ctx->P, ctx->cv, ts, qxs, qys,
                                   ctx->ssraw);
        }
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(integrals_kernel, dim3(nblocks(ctx->N, 256)), dim3(256), 0, ctx->stream, ctx->ssraw,
                           qxs, qys, ctx->int_tri, ctx->N, ctx->P, ctx->cv, temp, qx, qy);
        HIP_TRY(hipGetLastError());
        return 0;
    });
}

extern "C" int pbte_wall_flux(pbte_ctx* ctx, int32_t precision, const void* vdf, double* flux_wall)
{
    return guarded(ctx, [&]() {
        if (!ctx->n_wall) return 0;
#define WF(PP)                                                                                          \
    case PP:                                                                                            \
        if (precision == PBTE_FP64)                                                                     \
            hipLaunchKernelGGL((wall_flux_kernel<double, PP>), dim3(ctx->n_wall), dim3(64), 0, ctx->stream, \
                               (const double*)vdf, ctx->cxv, ctx->cyv, ctx->domega, ctx->fcm, ctx->nrm,  \
                               ctx->wall_face, ctx->wall_tri, ctx->wall_lfc, ctx->D, ctx->N, flux_wall); \
        else                                                                                            \
            hipLaunchKernelGGL((wall_flux_kernel<float, PP>), dim3(ctx->n_wall), dim3(64), 0, ctx->stream, \
                               (const float*)vdf, ctx->cxv, ctx->cyv, ctx->domega, ctx->fcm, ctx->nrm,   \
                               ctx->wall_face, ctx->wall_tri, ctx->wall_lfc, ctx->D, ctx->N, flux_wall); \
        break;
        switch (ctx->P) {
            WF(3) WF(6) WF(10) WF(15)
            default: throw std::string("wall_flux: unsupported P");
        }
#undef WF
        HIP_TRY(hipGetLastError());
        return 0;
    });
}

extern "C" int pbte_transport_residual(pbte_ctx* ctx, const double* ts, const double* qxs, const double* qys,
                                       const double* vdf, const double* flux_wall, double* out_host)
{
    return guarded(ctx, [&]() {
        if (ctx->n_wall && !flux_wall) throw std::string("adiabatic walls need flux_wall");
        launch_source<double>(ctx, ts, qxs, qys, ctx->src64, pad4(3 * ctx->P));
        const unsigned gx = nblocks(ctx->N, 256);
        double* part = ctx->red;
        double* out = ctx->red + (size_t)ctx->D * gx * 2;
        if ((size_t)ctx->D * gx * 2 + 2 * (size_t)ctx->D > ctx->red_len) throw std::string("reduction buffer too small");
        const dim3 cgrid(gx, (ctx->D + 63) / 64);
#define RS(PP) case PP: hipLaunchKernelGGL(residual_cells_kernel<PP>, cgrid, dim3(256), 0, ctx->stream, vdf, \
        ctx->src64, ctx->cxv, ctx->cyv, ctx->domega, ctx->mass, ctx->gx, ctx->gy, ctx->fcm, ctx->geo64, ctx->K64, \
        ctx->bvec64, flux_wall, 1.0 / ctx->tau_c, ctx->D, ctx->N, part); break;
        switch (ctx->P) { RS(3) RS(6) RS(10) RS(15) }
#undef RS
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(rowsum_kernel, dim3(ctx->D * 2), dim3(256), 0, ctx->stream, part, ctx->D, (int)gx, 2, 2,
                           out);
        HIP_TRY(hipGetLastError());
        double* h = ctx->host_staging((size_t)2 * ctx->D);
        HIP_TRY(hipMemcpyAsync(h, out, sizeof(double) * 2 * ctx->D, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        for (int d = 0; d < ctx->D; ++d) {
            out_host[d] = h[2 * d];
            out_host[ctx->D + d] = h[2 * d + 1];
        }
        return 0;
    });
}

// residual and mass into out[0], out[1] (device memory), stream-ordered
static void launch_iterate_residual(pbte_ctx* ctx, const double* temp, const double* t_old, double* out)
{
    const int nbk = (int)std::min<int64_t>(kIterBlocks, std::max<int64_t>(1, ((int64_t)ctx->N + 255) / 256));
    if ((size_t)3 * nbk > ctx->red_len) throw std::string("reduction buffer too small");
    hipLaunchKernelGGL(iterate_partial_kernel, dim3(nbk), dim3(256), 0, ctx->stream, temp, t_old, ctx->N, ctx->red);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(iterate_residual_kernel, dim3(1), dim3(256), 0, ctx->stream, ctx->red, nbk, out);
    HIP_TRY(hipGetLastError());
}

extern "C" int pbte_residual_iterate_async(pbte_ctx* ctx, const double* temp, const double* t_old, double* out_dev)
{
    return guarded(ctx, [&]() {
        if (!out_dev) throw std::string("residual_iterate_async needs an output");
        launch_iterate_residual(ctx, temp, t_old, out_dev);
        return 0;
    });
}

extern "C" int pbte_residual_iterate(pbte_ctx* ctx, const double* temp, const double* t_old, double* res_host,
                                     double* mass_host)
{
    return guarded(ctx, [&]() {
        launch_iterate_residual(ctx, temp, t_old, ctx->small);
        double* h = ctx->host_staging(2);
        HIP_TRY(hipMemcpyAsync(h, ctx->small, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        if (res_host) *res_host = h[0];
        if (mass_host) *mass_host = h[1];
        return 0;
    });
}

extern "C" int pbte_vec_multidot(pbte_ctx* ctx, const double* Q, int64_t ld, int32_t m, const double* w,
                                 int64_t n, double* c_host)
{
    return guarded(ctx, [&]() {
        if (m <= 0) return 0;
        if (m > kMaxBasis || (size_t)m * kDotBlocks + m > ctx->red_len) throw std::string("too many basis vectors");
        const int nbk = (int)std::min<int64_t>(kDotBlocks, std::max<int64_t>(1, (n + 255) / 256));
        dim3 grid(nbk, (m + kRows - 1) / kRows);
        hipLaunchKernelGGL(multidot_kernel, grid, dim3(256), 0, ctx->stream, Q, ld, m, w, n, ctx->red);
        HIP_TRY(hipGetLastError());
        double* out = ctx->red + (size_t)m * nbk;
        hipLaunchKernelGGL(rowsum_kernel, dim3(m), dim3(256), 0, ctx->stream, ctx->red, m, nbk, 1, 1, out);
        HIP_TRY(hipGetLastError());
        double* h = ctx->host_staging((size_t)m);
        HIP_TRY(hipMemcpyAsync(h, out, sizeof(double) * m, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        std::memcpy(c_host, h, sizeof(double) * m);
        return 0;
    });
}

extern "C" int pbte_vec_update(pbte_ctx* ctx, const double* Q, int64_t ld, int32_t m, const double* c_host,
                               double* w, int64_t n)
{
    return guarded(ctx, [&]() {
        if (m <= 0) return 0;
        if (m > kMaxBasis) throw std::string("too many basis vectors");
        double* h = ctx->host_staging((size_t)m);
        HIP_TRY(hipStreamSynchronize(ctx->stream));   // staging buffer may still feed a copy
        std::memcpy(h, c_host, sizeof(double) * m);
        HIP_TRY(hipMemcpyAsync(ctx->small, h, sizeof(double) * m, hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(update_kernel, dim3(nblocks(n, 256)), dim3(256), 0, ctx->stream, Q, ld, m, ctx->small,
                           w, n);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        return 0;
    });
}

// CGS2 into coef = (c1 [m], c2 [m], |w|^2) on the stream: multidot + rowsum
// for c1, then (m <= 32) update + multidot fused for c2 and update + norm
// fused for |w|^2 (two reads of Q instead of four), each followed by its
// rowsum; the same bits as the unfused multidot / update sequence
static void launch_cgs2(pbte_ctx* ctx, const double* Q, int64_t ld, int m, double* w, int64_t n, int nbk,
                        double* coef)
{
    constexpr int kFuseRows = 32;
    const dim3 grid(nbk, (m + kRows - 1) / kRows);
    double* c1 = coef;
    double* c2 = coef + m;
    hipLaunchKernelGGL(multidot_kernel, grid, dim3(256), 0, ctx->stream, Q, ld, m, w, n, ctx->red);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(rowsum_kernel, dim3(m), dim3(256), 0, ctx->stream, ctx->red, m, nbk, 1, 1, c1);
    HIP_TRY(hipGetLastError());
    if (m <= kFuseRows) {
        // accumulators for the rows present, in steps of eight
        auto* dot = m <= 8 ? update_dot_kernel<8, false> : m <= 16 ? update_dot_kernel<16, false>
                  : m <= 24 ? update_dot_kernel<24, false> : update_dot_kernel<kFuseRows, false>;
        hipLaunchKernelGGL(dot, dim3(nbk), dim3(256), 0, ctx->stream, Q, ld, m, c1, w, n, ctx->red);
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(rowsum_kernel, dim3(m), dim3(256), 0, ctx->stream, ctx->red, m, nbk, 1, 1, c2);
        HIP_TRY(hipGetLastError());
        auto* nrm = m <= 8 ? update_dot_kernel<8, true> : m <= 16 ? update_dot_kernel<16, true>
                  : m <= 24 ? update_dot_kernel<24, true> : update_dot_kernel<kFuseRows, true>;
        hipLaunchKernelGGL(nrm, dim3(nbk), dim3(256), 0, ctx->stream, Q, ld, m, c2, w, n, ctx->red);
        HIP_TRY(hipGetLastError());
    } else {
        hipLaunchKernelGGL(update_kernel, dim3(nblocks(n, 256)), dim3(256), 0, ctx->stream, Q, ld, m, c1, w, n);
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(multidot_kernel, grid, dim3(256), 0, ctx->stream, Q, ld, m, w, n, ctx->red);
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(rowsum_kernel, dim3(m), dim3(256), 0, ctx->stream, ctx->red, m, nbk, 1, 1, c2);
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(update_kernel, dim3(nblocks(n, 256)), dim3(256), 0, ctx->stream, Q, ld, m, c2, w, n);
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(multidot_kernel, dim3(nbk, 1), dim3(256), 0, ctx->stream, w, n, 1, w, n, ctx->red);
        HIP_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(rowsum_kernel, dim3(1), dim3(256), 0, ctx->stream, ctx->red, 1, nbk, 1, 1, coef + 2 * m);
    HIP_TRY(hipGetLastError());
}

// one classical Gram-Schmidt pass applied twice (krylov.py:203-207) on the
// device: c1 = Q w, w -= Q^T c1, c2 = Q w, w -= Q^T c2, then |w|^2; the
// coefficients never leave the device until the single copy of
// (c1, c2, |w|^2) at the end.  The same kernels and reduction order as
// pbte_vec_multidot / pbte_vec_update / a multidot of w with itself, so
// the results are those of that sequence bit for bit.
extern "C" int pbte_vec_cgs2(pbte_ctx* ctx, const double* Q, int64_t ld, int32_t m, double* w, int64_t n,
                             double* out_host)
{
    return guarded(ctx, [&]() {
        if (m <= 0) throw std::string("cgs2 needs at least one basis vector");
        if (m > kMaxBasis) throw std::string("too many basis vectors");
        const int nbk = (int)std::min<int64_t>(kDotBlocks, std::max<int64_t>(1, (n + 255) / 256));
        if ((size_t)m * nbk + 2 * (size_t)m + 1 > ctx->red_len) throw std::string("too many basis vectors");
        double* coef = ctx->red + (size_t)m * nbk;          // c1 [m], c2 [m], |w|^2
        launch_cgs2(ctx, Q, ld, m, w, n, nbk, coef);
        double* h = ctx->host_staging((size_t)2 * m + 1);
        HIP_TRY(hipMemcpyAsync(h, coef, sizeof(double) * (2 * (size_t)m + 1), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        std::memcpy(out_host, h, sizeof(double) * (2 * (size_t)m + 1));
        return 0;
    });
}

// The same Gram-Schmidt, returning at once: the coefficients are copied to
// pinned host memory behind an event, and (next != NULL) w / |w| is written
// into next on the device, so the caller can queue the next product before
// it waits for the coefficients in pbte_vec_cgs2_end.
extern "C" int pbte_vec_cgs2_begin(pbte_ctx* ctx, const double* Q, int64_t ld, int32_t m, double* w, int64_t n,
                                   double* next)
{
    return guarded(ctx, [&]() {
        if (m <= 0) throw std::string("cgs2 needs at least one basis vector");
        if (m > kMaxBasis) throw std::string("too many basis vectors");
        if (ctx->cg_m) {
            // a pair left open (the caller raised between begin and end):
            // let its copy land, then drop it
            HIP_TRY(hipEventSynchronize(ctx->cg_event));
            ctx->cg_m = 0;
        }
        const int nbk = (int)std::min<int64_t>(kDotBlocks, std::max<int64_t>(1, (n + 255) / 256));
        if ((size_t)m * nbk + 2 * (size_t)m + 1 > ctx->red_len) throw std::string("too many basis vectors");
        if (!ctx->cg_host) {
            HIP_TRY(hipHostMalloc((void**)&ctx->cg_host, sizeof(double) * (2 * kMaxBasis + 1), hipHostMallocDefault));
            HIP_TRY(hipEventCreateWithFlags(&ctx->cg_event, hipEventDisableTiming));
        }
        double* coef = ctx->red + (size_t)m * nbk;
        launch_cgs2(ctx, Q, ld, m, w, n, nbk, coef);
        HIP_TRY(hipMemcpyAsync(ctx->cg_host, coef, sizeof(double) * (2 * (size_t)m + 1), hipMemcpyDeviceToHost,
                               ctx->stream));
        HIP_TRY(hipEventRecord(ctx->cg_event, ctx->stream));
        if (next) {
            hipLaunchKernelGGL(normalize_kernel, dim3(nblocks(n, 256)), dim3(256), 0, ctx->stream, coef + 2 * m, w,
                               next, n);
            HIP_TRY(hipGetLastError());
        }
        ctx->cg_m = m;
        return 0;
    });
}

extern "C" int pbte_vec_cgs2_end(pbte_ctx* ctx, double* out_host)
{
    return guarded(ctx, [&]() {
        if (!ctx->cg_m) throw std::string("cgs2_end without cgs2_begin");
        const int m = ctx->cg_m;
        ctx->cg_m = 0;
        HIP_TRY(hipEventSynchronize(ctx->cg_event));
        std::memcpy(out_host, ctx->cg_host, sizeof(double) * (2 * (size_t)m + 1));
        return 0;
    });
}

// out = (v - t) + g (g may be NULL: out = v - t), the fixed-point defect of
// a product in the same rounding as a copy and two axpby calls
extern "C" int pbte_vec_defect(pbte_ctx* ctx, const double* v, const double* t, const double* g, double* out,
                               int64_t n)
{
    return guarded(ctx, [&]() {
        hipLaunchKernelGGL(fixed_point_defect_kernel, dim3(nblocks(n, 256)), dim3(256), 0, ctx->stream, v, t, g, out,
                           n);
        HIP_TRY(hipGetLastError());
        return 0;
    });
}

// step_host = max |o - d| (NaN if any entry is NaN), then d += omega (o - d)
// unless omega == 0; waits for the stream
extern "C" int pbte_vec_relax(pbte_ctx* ctx, double omega, const double* o, double* d, int64_t n,
                              double* step_host)
{
    return guarded(ctx, [&]() {
        if (n <= 0) {
            *step_host = 0.0;
            return 0;
        }
        const int nbk = (int)std::min<int64_t>(kDotBlocks, std::max<int64_t>(1, (n + 255) / 256));
        if ((size_t)nbk + 1 > ctx->red_len) throw std::string("reduction buffer too small");
        hipLaunchKernelGGL(relax_kernel, dim3(nbk), dim3(256), 0, ctx->stream, omega, o, d, n, ctx->red);
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(max_final_kernel, dim3(1), dim3(256), 0, ctx->stream, ctx->red, nbk, ctx->red + nbk);
        HIP_TRY(hipGetLastError());
        double* h = ctx->host_staging(1);
        HIP_TRY(hipMemcpyAsync(h, ctx->red + nbk, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        *step_host = h[0];
        return 0;
    });
}

extern "C" int pbte_vec_axpby(pbte_ctx* ctx, double a, const double* x, double b, double* y, int64_t n)
{
    return guarded(ctx, [&]() {
        hipLaunchKernelGGL(axpby_kernel, dim3(nblocks(n, 256)), dim3(256), 0, ctx->stream, a, x, b, y, n);
        HIP_TRY(hipGetLastError());
        return 0;
    });
}

extern "C" int pbte_periodic_gather(pbte_ctx* ctx, int32_t precision, const void* vdf, double* out)
{
    return guarded(ctx, [&]() {
        const int64_t n = (int64_t)ctx->D * ctx->n_per * ctx->P;
        if (n == 0) return 0;
        if (precision == PBTE_FP64)
            hipLaunchKernelGGL(periodic_gather_kernel<double>, dim3(nblocks(n, 256)), dim3(256), 0, ctx->stream,
                               (const double*)vdf, ctx->per_tri, ctx->n_per, ctx->D, ctx->N, ctx->P, out);
        else
            hipLaunchKernelGGL(periodic_gather_kernel<float>, dim3(nblocks(n, 256)), dim3(256), 0, ctx->stream,
                               (const float*)vdf, ctx->per_tri, ctx->n_per, ctx->D, ctx->N, ctx->P, out);
        HIP_TRY(hipGetLastError());
        return 0;
    });
}

extern "C" int pbte_periodic_scatter(pbte_ctx* ctx, int32_t precision, const double* in, void* pbuf)
{
    return guarded(ctx, [&]() {
        const int64_t n = (int64_t)ctx->D * ctx->n_per * ctx->P;
        if (n == 0) return 0;
        if (precision == PBTE_FP64)
            hipLaunchKernelGGL(periodic_scatter_kernel<double>, dim3(nblocks(n, 256)), dim3(256), 0, ctx->stream,
                               in, ctx->per_face, ctx->n_per, ctx->D, ctx->F, ctx->P, (double*)pbuf);
        else
            hipLaunchKernelGGL(periodic_scatter_kernel<float>, dim3(nblocks(n, 256)), dim3(256), 0, ctx->stream,
                               in, ctx->per_face, ctx->n_per, ctx->D, ctx->F, ctx->P, (float*)pbuf);
        HIP_TRY(hipGetLastError());
        return 0;
    });
}

extern "C" int pbte_gsis_setup(pbte_ctx* ctx, const pbte_gsis_data* g)
{
    return guarded(ctx, [&]() {
        const int P = ctx->P, Q = ctx->Q, N = ctx->N, F = ctx->F, D = ctx->D;
        std::vector<double> hw((size_t)4 * D);
        for (int d = 0; d < D; ++d) {
            const double cx = g->cxv[d], cy = g->cyv[d], w = g->domega[d];
            hw[d] = cx * (5.0 * cx * cx - 3.0) * w;
            hw[D + d] = cy * (5.0 * cx * cx - 1.0) * w;
            hw[2 * D + d] = cx * (5.0 * cy * cy - 1.0) * w;
            hw[3 * D + d] = cy * (5.0 * cy * cy - 3.0) * w;
        }
        ctx->gs_hw = ctx->upload(hw.data(), hw.size());
        // element operators from the device inverse, AA_TRACE and BA_SOL
        const int K = 3 * P + 9 * Q;
        ctx->gs_op = ctx->alloc<double>((size_t)N * 4 * P * K);
        hipLaunchKernelGGL(gsis_op_kernel, dim3(nblocks((int64_t)N * 4 * P * K, 256)), dim3(256), 0, ctx->stream,
                           g->inv_aa_sol, g->ba_sol, (int64_t)N, P, Q, 3 * P, 4 * P, ctx->gs_op);
        ctx->gs_op_hi = ctx->alloc<double>((size_t)N * 4 * P * 4 * P);
        hipLaunchKernelGGL(inv_rows_kernel, dim3(nblocks((int64_t)N * 16 * P * P, 256)), dim3(256), 0, ctx->stream,
                           g->inv_aa_sol, (int64_t)N, P, 3 * P, 4 * P, ctx->gs_op_hi);
        ctx->gs_aat = ctx->alloc<double>((size_t)N * 9 * Q * 3 * P);
        hipLaunchKernelGGL(aat_rows_kernel, dim3(nblocks((int64_t)N * 27 * Q * P, 256)), dim3(256), 0, ctx->stream,
                           g->aa_trace, (int64_t)N, P, Q, 0, 3 * P, ctx->gs_aat);
        ctx->gs_aat_hi = ctx->alloc<double>((size_t)N * 9 * Q * 4 * P);
        hipLaunchKernelGGL(aat_rows_kernel, dim3(nblocks((int64_t)N * 36 * Q * P, 256)), dim3(256), 0, ctx->stream,
                           g->aa_trace, (int64_t)N, P, Q, 3 * P, 4 * P, ctx->gs_aat_hi);
        HIP_TRY(hipGetLastError());
        ctx->gs_side = ctx->alloc<double>((size_t)N * 3 * 3 * Q);
        HIP_TRY(hipGetLastError());
        ctx->gs_t1 = ctx->upload(g->t1, F);
        ctx->gs_l1 = ctx->upload(g->l1, F);
        ctx->gs_t2 = ctx->upload(g->t2, F);
        ctx->gs_l2 = ctx->upload(g->l2, F);
        ctx->gs_flip = ctx->upload(g->flip, F);
        ctx->gs_nw = g->n_wall;
        ctx->gs_nth = g->n_wall_th;
        ctx->gs_wface = ctx->upload(g->wall_face, std::max(g->n_wall, 1));
        ctx->gs_wtri = ctx->upload(g->wall_tri, std::max(g->n_wall, 1));
        ctx->gs_wnx = ctx->upload(g->wall_nx, std::max(g->n_wall, 1));
        ctx->gs_wny = ctx->upload(g->wall_ny, std::max(g->n_wall, 1));
        // wall face slices of int_tri_fc_fc [P][Q][N][3] -> [k][Q][P]
        std::vector<double> wt((size_t)std::max(g->n_wall, 1) * Q * P, 0.0);
        for (int k = 0; k < g->n_wall; ++k)
            for (int m = 0; m < Q; ++m)
                for (int l = 0; l < P; ++l)
                    wt[((size_t)k * Q + m) * P + l] =
                        g->tri_fc_fc[(((size_t)l * Q + m) * N + g->wall_tri[k]) * 3 + g->wall_lfc[k]];
        ctx->gs_wtfc = ctx->upload(wt.data(), wt.size());
        ctx->gs_hmin = ctx->upload(g->tri_hmin, N);
        ctx->gs_tri_faces = ctx->upload(g->tri_faces, (size_t)N * 3);
        ctx->gs_wmom = ctx->alloc<double>((size_t)std::max(g->n_wall, 1) * 3 * Q);
        ctx->gs_tc5 = g->tc5;
        ctx->gs_tau_thr = g->tau_thr;
        ctx->gs_drift = g->drift;
        ctx->gs_variant_b = g->variant_b;
        ctx->gs_use_hmin = g->use_hmin;
        ctx->gsis = true;
        return 0;
    });
}

extern "C" int pbte_gsis_moments(pbte_ctx* ctx, int32_t precision, const void* vdf, double* mom7)
{
    return guarded(ctx, [&]() {
        if (!ctx->gsis) throw std::string("pbte_gsis_setup first");
        const int64_t np = (int64_t)ctx->N * ctx->P;
        if (precision == PBTE_FP64) {
            if (np % 2 == 0)
                hipLaunchKernelGGL((moments7_kernel<double, 2>), dim3(nblocks(np / 2, 256)), dim3(256), 0,
                                   ctx->stream, (const double*)vdf, ctx->cxv, ctx->cyv, ctx->domega, ctx->gs_hw,
                                   ctx->D, ctx->N, ctx->P, mom7);
            else
                hipLaunchKernelGGL((moments7_kernel<double, 1>), dim3(nblocks(np, 256)), dim3(256), 0, ctx->stream,
                                   (const double*)vdf, ctx->cxv, ctx->cyv, ctx->domega, ctx->gs_hw, ctx->D, ctx->N,
                                   ctx->P, mom7);
        } else {
            if (np % 4 == 0)
                hipLaunchKernelGGL((moments7_kernel<float, 4>), dim3(nblocks(np / 4, 256)), dim3(256), 0,
                                   ctx->stream, (const float*)vdf, ctx->cxv, ctx->cyv, ctx->domega, ctx->gs_hw,
                                   ctx->D, ctx->N, ctx->P, mom7);
            else
                hipLaunchKernelGGL((moments7_kernel<float, 1>), dim3(nblocks(np, 256)), dim3(256), 0, ctx->stream,
                                   (const float*)vdf, ctx->cxv, ctx->cyv, ctx->domega, ctx->gs_hw, ctx->D, ctx->N,
                                   ctx->P, mom7);
        }
        HIP_TRY(hipGetLastError());
        return 0;
    });
}

// the binary32 copies of the step's operators, made once
static void gsis_narrow(pbte_ctx* ctx)
{
    if (ctx->gs_op32) return;
    const int64_t N = ctx->N, P = ctx->P, Q = ctx->Q, K = 3 * P + 9 * Q;
    const int64_t nop = N * 4 * P * K, naat = N * 9 * Q * 3 * P;
    ctx->gs_op32 = ctx->alloc<float>((size_t)nop);
    ctx->gs_aat32 = ctx->alloc<float>((size_t)naat);
    hipLaunchKernelGGL(narrow_kernel<float>, dim3(4096), dim3(256), 0, ctx->stream, ctx->gs_op, ctx->gs_op32, nop);
    hipLaunchKernelGGL(narrow_kernel<float>, dim3(4096), dim3(256), 0, ctx->stream, ctx->gs_aat, ctx->gs_aat32, naat);
    HIP_TRY(hipGetLastError());
}

extern "C" int pbte_gsis_source(pbte_ctx* ctx, int32_t precision, const double* mom7, double* src)
{
    return guarded(ctx, [&]() {
        if (!ctx->gsis) throw std::string("pbte_gsis_setup first");
        if (precision != PBTE_FP64 && precision != PBTE_FP32) throw std::string("gsis_source: bad precision");
        const int64_t N = ctx->N;
        if (precision == PBTE_FP32) gsis_narrow(ctx);
        const dim3 grid((unsigned)((N + 7) / 8));
#define HS(PP, QQ)                                                                                            \
    case PP:                                                                                                  \
        if (precision == PBTE_FP64)                                                                           \
            hipLaunchKernelGGL((hot_side_kernel<PP, QQ, double>), grid, dim3(256), 0, ctx->stream, mom7, ctx->gx, \
                               ctx->gy, ctx->gs_op, N, ctx->gs_tc5, src, ctx->gs_side);                       \
        else                                                                                                  \
            hipLaunchKernelGGL((hot_side_kernel<PP, QQ, float>), grid, dim3(256), 0, ctx->stream, mom7, ctx->gx, \
                               ctx->gy, ctx->gs_op32, N, ctx->gs_tc5, src, ctx->gs_side);                     \
        break;
        switch (ctx->P) {
            HS(3, 2) HS(6, 3) HS(10, 4) HS(15, 5)
            default: throw std::string("gsis_source: unsupported P");
        }
#undef HS
        HIP_TRY(hipGetLastError());
        return 0;
    });
}

extern "C" int pbte_gsis_rhs(pbte_ctx* ctx, int32_t precision, const void* vdf, double* ffa)
{
    return guarded(ctx, [&]() {
        if (!ctx->gsis) throw std::string("pbte_gsis_setup first");
        const int N = ctx->N, P = ctx->P, Q = ctx->Q, F = ctx->F;
        hipLaunchKernelGGL(trace_rhs_kernel, dim3(nblocks((int64_t)F * 3 * Q, 256)), dim3(256), 0, ctx->stream,
                           ctx->gs_side, ctx->gs_t1, ctx->gs_l1, ctx->gs_t2, ctx->gs_l2, ctx->gs_flip, F, Q, ffa);
        HIP_TRY(hipGetLastError());
        if (ctx->gs_nw) {
            const unsigned g = nblocks((int64_t)ctx->gs_nw * Q, 256);
            if (!vdf)
                hipLaunchKernelGGL(fill_kernel, dim3(nblocks((int64_t)ctx->gs_nw * 3 * Q, 256)), dim3(256), 0,
                                   ctx->stream, ctx->gs_wmom, (int64_t)ctx->gs_nw * 3 * Q, 0.0);
            else {
#define WT3(PP, QQ)                                                                                        \
    case PP:                                                                                               \
        if (precision == PBTE_FP64)                                                                        \
            hipLaunchKernelGGL((wall_trace_kernel<double, PP, QQ>), dim3(ctx->gs_nw), dim3(64), 0, ctx->stream, \
                               (const double*)vdf, ctx->cxv, ctx->cyv, ctx->domega, ctx->gs_wtfc, ctx->gs_wtri, \
                               ctx->D, N, ctx->gs_wmom);                                                   \
        else                                                                                               \
            hipLaunchKernelGGL((wall_trace_kernel<float, PP, QQ>), dim3(ctx->gs_nw), dim3(64), 0, ctx->stream, \
                               (const float*)vdf, ctx->cxv, ctx->cyv, ctx->domega, ctx->gs_wtfc, ctx->gs_wtri, \
                               ctx->D, N, ctx->gs_wmom);                                                   \
        break;
                switch (P) {
                    WT3(3, 2) WT3(6, 3) WT3(10, 4) WT3(15, 5)
                    default: throw std::string("gsis_rhs: unsupported P");
                }
#undef WT3
            }
            HIP_TRY(hipGetLastError());
            hipLaunchKernelGGL(trace_wall_kernel, dim3(g), dim3(256), 0, ctx->stream, ctx->gs_wmom, ctx->gs_wface,
                               ctx->gs_wnx, ctx->gs_wny, ctx->gs_nw, ctx->gs_nth, Q, ctx->cv, ctx->gs_variant_b, ffa);
            HIP_TRY(hipGetLastError());
        }
        return 0;
    });
}

extern "C" int pbte_gsis_precond_setup(pbte_ctx* ctx, const double* inv_aa_sol, const double* ba_sol)
{
    return guarded(ctx, [&]() {
        if (!ctx->gsis) throw std::string("pbte_gsis_setup first");
        const int P = ctx->P, Q = ctx->Q;
        const int64_t N = ctx->N, K = 3 * P + 9 * Q;
        if (!ctx->gs_op_pc) ctx->gs_op_pc = ctx->alloc<double>((size_t)N * 3 * P * K);
        hipLaunchKernelGGL(gsis_op_kernel, dim3(nblocks(N * 3 * P * K, 256)), dim3(256), 0, ctx->stream, inv_aa_sol,
                           ba_sol, N, P, Q, 0, 3 * P, ctx->gs_op_pc);
        HIP_TRY(hipGetLastError());
        return 0;
    });
}

extern "C" int pbte_gsis_precond_src(pbte_ctx* ctx, const double* r, double sign, double* src)
{
    return guarded(ctx, [&]() {
        if (!ctx->gs_op_pc) throw std::string("pbte_gsis_precond_setup first");
        const int64_t N = ctx->N;
        hipLaunchKernelGGL(precond_side_kernel<double>, dim3((unsigned)((N + 7) / 8)), dim3(256), 0, ctx->stream, r,
                           sign, ctx->gs_op_pc, N, ctx->P, ctx->Q, src, ctx->gs_side);
        HIP_TRY(hipGetLastError());
        return 0;
    });
}

extern "C" int pbte_gsis_local(pbte_ctx* ctx, int32_t precision, const double* src, const double* u_trace,
                               double* uq)
{
    return guarded(ctx, [&]() {
        if (!ctx->gsis) throw std::string("pbte_gsis_setup first");
        const int64_t N = ctx->N, P = ctx->P;
        const unsigned g = nblocks(N * 3 * P, 256);
        if (precision == PBTE_FP64) {
            hipLaunchKernelGGL(local3_kernel<double>, dim3(g), dim3(256), 0, ctx->stream, ctx->gs_aat, src, u_trace,
                               ctx->gs_tri_faces, N, ctx->P, ctx->Q, uq);
        } else if (precision == PBTE_FP32) {
            gsis_narrow(ctx);
            hipLaunchKernelGGL(local3_kernel<float>, dim3(g), dim3(256), 0, ctx->stream, ctx->gs_aat32, src, u_trace,
                               ctx->gs_tri_faces, N, ctx->P, ctx->Q, uq);
        } else {
            throw std::string("gsis_local: bad precision");
        }
        HIP_TRY(hipGetLastError());
        return 0;
    });
}

extern "C" int pbte_gsis_uq(pbte_ctx* ctx, const double* mom7, const double* u_trace, double* uq)
{
    return guarded(ctx, [&]() {
        if (!ctx->gsis) throw std::string("pbte_gsis_setup first");
        const int64_t N = ctx->N;
#define UQ(PP, QQ)                                                                                           \
    case PP:                                                                                                 \
        hipLaunchKernelGGL((uq_full_kernel<PP, QQ>), dim3((unsigned)((N + 3) / 4)), dim3(256), 0, ctx->stream, \
                           mom7, ctx->gx, ctx->gy, ctx->gs_op, ctx->gs_op_hi, ctx->gs_aat, ctx->gs_aat_hi,       \
                           u_trace, ctx->gs_tri_faces, N, ctx->gs_tc5, uq);                                  \
        break;
        switch (ctx->P) {
            UQ(3, 2) UQ(6, 3) UQ(10, 4) UQ(15, 5)
            default: throw std::string("gsis_uq: unsupported P");
        }
#undef UQ
        HIP_TRY(hipGetLastError());
        return 0;
    });
}

extern "C" int pbte_gsis_correct(pbte_ctx* ctx, int32_t precision, void* vdf, const double* mom7, const double* uq,
                                 const double* defect, double* ts, double* qxs, double* qys, double* temp,
                                 double* qx, double* qy, double* mkin)
{
    return guarded(ctx, [&]() {
        if (!ctx->gsis) throw std::string("pbte_gsis_setup first");
        const int N = ctx->N, P = ctx->P;
        const unsigned g = nblocks((int64_t)N * P, 256);
        if (precision == PBTE_FP64)
            hipLaunchKernelGGL(correct_kernel<double>, dim3(g), dim3(256), 0, ctx->stream, (double*)vdf, mom7, uq,
                               defect, ctx->cxv, ctx->cyv, ctx->gs_hmin, ctx->D, N, P, ctx->cv, ctx->pi, ctx->tau_r,
                               ctx->gs_tau_thr, ctx->gs_drift, ctx->gs_use_hmin, ts, qxs, qys, mkin);
        else
            hipLaunchKernelGGL(correct_kernel<float>, dim3(g), dim3(256), 0, ctx->stream, (float*)vdf, mom7, uq,
                               defect, ctx->cxv, ctx->cyv, ctx->gs_hmin, ctx->D, N, P, ctx->cv, ctx->pi, ctx->tau_r,
                               ctx->gs_tau_thr, ctx->gs_drift, ctx->gs_use_hmin, ts, qxs, qys, mkin);
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(dof_integrals_kernel, dim3(nblocks(N, 256)), dim3(256), 0, ctx->stream, ts, qxs, qys,
                           ctx->int_tri, N, P, temp, qx, qy);
        HIP_TRY(hipGetLastError());
        return 0;
    });
}

extern "C" int pbte_gsis_correct_moments(pbte_ctx* ctx, const double* mom7, const double* uq, const double* defect,
                                         double* ts, double* qxs, double* qys, double* temp, double* qx, double* qy,
                                         double* mkin, double* coef)
{
    return guarded(ctx, [&]() {
        if (!ctx->gsis) throw std::string("pbte_gsis_setup first");
        const int N = ctx->N, P = ctx->P;
        hipLaunchKernelGGL(correct_moments_kernel, dim3(nblocks((int64_t)N * P, 256)), dim3(256), 0, ctx->stream,
                           mom7, uq, defect, ctx->gs_hmin, N, P, ctx->cv, ctx->pi, ctx->tau_r, ctx->gs_tau_thr,
                           ctx->gs_use_hmin, ts, qxs, qys, mkin, coef);
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(dof_integrals_kernel, dim3(nblocks(N, 256)), dim3(256), 0, ctx->stream, ts, qxs, qys,
                           ctx->int_tri, N, P, temp, qx, qy);
        HIP_TRY(hipGetLastError());
        return 0;
    });
}

extern "C" int pbte_gsis_apply_correction(pbte_ctx* ctx, int32_t precision, void* vdf, const double* coef,
                                          const int32_t* cells, int32_t ncells)
{
    return guarded(ctx, [&]() {
        if (!ctx->gsis) throw std::string("pbte_gsis_setup first");
        const int N = ctx->N, P = ctx->P, D = ctx->D;
        const int64_t np = (int64_t)N * P;
        const double dr = ctx->gs_drift;
        if (cells) {
            if (ncells <= 0) return 0;
            const unsigned g = nblocks((int64_t)ncells * P * D, 256);
            if (precision == PBTE_FP64)
                hipLaunchKernelGGL(apply_correction_cells_kernel<double>, dim3(g), dim3(256), 0, ctx->stream,
                                   (double*)vdf, coef, ctx->cxv, ctx->cyv, cells, ncells, D, N, P, dr);
            else
                hipLaunchKernelGGL(apply_correction_cells_kernel<float>, dim3(g), dim3(256), 0, ctx->stream,
                                   (float*)vdf, coef, ctx->cxv, ctx->cyv, cells, ncells, D, N, P, dr);
        } else if (precision == PBTE_FP64) {
            if (np % 2 == 0)
                hipLaunchKernelGGL((apply_correction_kernel<double, 2>), dim3(nblocks(np / 2, 256)), dim3(256), 0,
                                   ctx->stream, (double*)vdf, coef, ctx->cxv, ctx->cyv, nullptr, 0, D, N, P, dr);
            else
                hipLaunchKernelGGL((apply_correction_kernel<double, 1>), dim3(nblocks(np, 256)), dim3(256), 0,
                                   ctx->stream, (double*)vdf, coef, ctx->cxv, ctx->cyv, nullptr, 0, D, N, P, dr);
        } else {
            if (np % 4 == 0)
                hipLaunchKernelGGL((apply_correction_kernel<float, 4>), dim3(nblocks(np / 4, 256)), dim3(256), 0,
                                   ctx->stream, (float*)vdf, coef, ctx->cxv, ctx->cyv, nullptr, 0, D, N, P, dr);
            else
                hipLaunchKernelGGL((apply_correction_kernel<float, 1>), dim3(nblocks(np, 256)), dim3(256), 0,
                                   ctx->stream, (float*)vdf, coef, ctx->cxv, ctx->cyv, nullptr, 0, D, N, P, dr);
        }
        HIP_TRY(hipGetLastError());
        return 0;
    });
}
