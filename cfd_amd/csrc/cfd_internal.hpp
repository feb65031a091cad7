// Internal declarations shared by the gfx950 kernels and the C-ABI layer.
//
// Layout in HBM (one copy per mesh / session, all device-resident):
//   nodal   U, U1, RHS          double4 per node (the reference's U(4,npoin))
//           T, VEL_X, ... , M   one double per node
//   element inpoel              3 x int32 per element, 1-based node ids
//           dNx, dNy            3 doubles per element (dNx(3,nelem))
//           area, dtl, shoc ... one double per element
//   tables  esup1/esup2, psup1/psup2, lap_rowptr/lap_idx as in the reference
//           (1-based ids, 0-based offsets), plus the slot maps
//           epos(3,nelem): index in esup1 of element e in node inpoel(k,e)'s list
//           lslot(3, 3*nelem): CSR position of (node of esup entry s, inpoel(j,e))
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cfd_abi.h"

namespace cfd {

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void set_last_error(const std::string& msg);

#define CFD_HIP(call)                                                                  \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            throw ::cfd::Error(std::string(#call) + ": " + hipGetErrorString(e_) +      \
                               " (" __FILE__ ":" + std::to_string(__LINE__) + ")");    \
    } while (0)

#define CFD_CHECK(cond, msg)                                                            \
    do {                                                                               \
        if (!(cond)) throw ::cfd::Error(msg);                                          \
    } while (0)

// C-ABI entry bodies: a cfd::Error becomes a status code by its message's
// prefix ("mesh:", "state:", "arg:", anything else a device error), any
// other exception a device error; the message is kept for cfd_last_error().
#define CFD_TRY(...)                                    \
    try {                                               \
        __VA_ARGS__;                                    \
        return CFD_OK;                                  \
    } catch (const cfd::Error& e) {                     \
        set_last_error(e.what());                       \
        std::string w = e.what();                       \
        if (w.find("mesh:") == 0) return CFD_ERR_MESH;  \
        if (w.find("state:") == 0) return CFD_ERR_STATE;\
        if (w.find("arg:") == 0) return CFD_ERR_ARG;    \
        return CFD_ERR_DEVICE;                          \
    } catch (const std::exception& e) {                 \
        set_last_error(e.what());                       \
        return CFD_ERR_DEVICE;                          \
    }

#define CFD_ARG(cond, msg) CFD_CHECK(cond, std::string("arg: ") + msg)

// Owned device allocation, or (borrow) a view of memory owned elsewhere,
// which release() leaves alone.
template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    bool owned = true;
    DBuf() = default;
    explicit DBuf(size_t count) { alloc(count); }
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : p(o.p), n(o.n), owned(o.owned) { o.p = nullptr; o.n = 0; }
    DBuf& operator=(DBuf&& o) noexcept {
        if (this != &o) {
            release();
            p = o.p; n = o.n; owned = o.owned;
            o.p = nullptr; o.n = 0;
        }
        return *this;
    }
    ~DBuf() { release(); }
    void alloc(size_t count) {
        release();
        n = count;
        owned = true;
        if (count) CFD_HIP(hipMalloc(&p, sizeof(T) * count));
    }
    void borrow(T* ptr, size_t count) {
        release();
        p = ptr;
        n = count;
        owned = false;
    }
    void release() {
        if (p && owned) (void)hipFree(p);
        p = nullptr;
        n = 0;
        owned = true;
    }
    void upload(const T* h, size_t count, hipStream_t s) {
        CFD_CHECK(count <= n, "upload larger than buffer");
        if (count) CFD_HIP(hipMemcpyAsync(p, h, sizeof(T) * count, hipMemcpyHostToDevice, s));
    }
    void download(T* h, size_t count, hipStream_t s) const {
        CFD_CHECK(count <= n, "download larger than buffer");
        if (count) CFD_HIP(hipMemcpyAsync(h, p, sizeof(T) * count, hipMemcpyDeviceToHost, s));
    }
    void zero(hipStream_t s) {
        if (n) CFD_HIP(hipMemsetAsync(p, 0, sizeof(T) * n, s));
    }
};

// Owned pinned host allocation (device-to-host reads without staging).
template <class T>
struct HostPinned {
    T* p = nullptr;
    HostPinned() = default;
    HostPinned(const HostPinned&) = delete;
    HostPinned& operator=(const HostPinned&) = delete;
    ~HostPinned() {
        if (p) (void)hipHostFree(p);
    }
    void alloc(size_t count) {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        CFD_HIP(hipHostMalloc(reinterpret_cast<void**>(&p), sizeof(T) * count));
    }
};

constexpr int kBlock = 256;
inline unsigned grid_for(size_t n, int block = kBlock) {
    size_t g = (n + block - 1) / block;
    return (unsigned)(g ? g : 1);
}

// Workgroups are dispatched to the 8 XCDs in turn (block b to XCD b mod 8),
// and each XCD has its own L2.  xcd_block(b, n) gives block b the logical
// block that makes XCD x work through the x-th eighth of [0, n) in order, so
// neighbouring logical blocks (neighbouring patches, neighbouring row
// blocks) mostly share an L2.  A bijection of [0, n); work and partial
// sums indexed by the logical block keep every result unchanged.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block(int b, int n) {
    const int q = n / kXcds, r = n % kXcds, x = b % kXcds, j = b / kXcds;
    return x * q + (x < r ? x : r) + j;
}

// ---------------------------------------------------------------------------
// Mesh tables (built on the device by tables.hip)
// ---------------------------------------------------------------------------
struct Tables {
    int npoin = 0, nelem = 0, npsup = 0, nnz = 0;
    DBuf<int> inpoel;            // 3*nelem
    DBuf<int> esup1, esup2;      // 3*nelem, npoin+1
    DBuf<int> epos;              // 3*nelem
    DBuf<int> psup1, psup2;      // npsup, npoin+1
    DBuf<int> rowptr, idx;       // npoin+1, nnz
    DBuf<int> lslot;             // 9*nelem (3 per esup entry)
};

void build_tables(Tables& t, const int* inpoel_host, int nelem, int npoin, hipStream_t s);
// Reorders each node's element list (esup1) by key[element] ascending and
// rebuilds the maps that follow that order (epos, lslot): a renumbered mesh
// then sums per node (MASAS, laplace, GCL) in the original element order.
void reorder_esup(Tables& t, const std::vector<int>& key, hipStream_t s);

// ---------------------------------------------------------------------------
// Kernel launchers (geometry, flow, PCG). All take device pointers and run on
// stream s.  Names follow the reference routine each one implements.
// ---------------------------------------------------------------------------
struct Consts {
    double gama, fr, fmu, fk, fcv, t_inf, rho_inf, u_inf, v_inf, cte, fsafe;
};

// deriv (subrutinas.f90:88-126); hmin_out: device double; partials holds
// grid_for(nelem) doubles
void launch_deriv(const Tables& t, const double* X, const double* Y, double* area, double* dNx,
                  double* dNy, double* HH, double* HHX, double* HHY, double* hmin_out,
                  double* partials, hipStream_t s);
// MASAS (subrutinas.f90:128-153) as an ascending-element gather
void launch_masas(const Tables& t, const double* area, double* M, hipStream_t s);
// mLaplace.f90:7-58 values (structure comes from Tables)
void launch_laplace(const Tables& t, const double* X, const double* Y, const double* dNx,
                    const double* dNy, double* lap, double* diag, hipStream_t s);
// gcl.f90:8-46 main; mode 0 = as written, 1 = corrected y-divergence; dt is a
// device scalar
void launch_gcl(const Tables& t, double* M, const double* Wx, const double* Wy,
                const double* Wxo, const double* Wyo, const double* dNx, const double* dNy,
                const double* area, const double* area_old, const double* dt, int mode,
                hipStream_t s);

// deltat (subrutinas.f90:155-218): DT and dtmin (device), clipped at 10*dtmin
void launch_deltat(const Tables& t, const Consts& c, const double* area, const double* T,
                   const double* VX, const double* VY, const double* WX, const double* WY,
                   double* DT, double* dtmin_out, double* partials, hipStream_t s);
// The element pass of deltat alone: unclipped DT and one min partial per
// block in partials; returns the number of partials (at most kRedBlocks).
constexpr int kRedBlocks = 2048;
int launch_deltat_partials(const Tables& t, const Consts& c, const double* area, const double* T,
                           const double* VX, const double* VY, const double* WX, const double* WY,
                           double* DT, double* partials, hipStream_t s);
// The same in patch order (one block per patch of ps, its node values in
// LDS): the same DT bits, one min partial per patch; returns ps.npatch
struct PatchSet;
int launch_deltat_patch(const PatchSet& ps, const Consts& c, const double* area, const double* T,
                        const double* VX, const double* VY, const double* WX, const double* WY,
                        double* DT, double* partials, hipStream_t s);
// The same with block b over the elements [chunk[b], chunk[b+1]): one min
// partial per chunk (batches keep each chunk inside one case)
void launch_deltat_chunks(const Tables& t, const Consts& c, const double* area, const double* T,
                          const double* VX, const double* VY, const double* WX, const double* WY,
                          double* DT, const int* chunk, int nchunks, double* partials,
                          const int* chunk_case, const Consts* cases, hipStream_t s);
// Inputs of the patch element residual (patch order, of the element type R:
// float on the FP32 path, double on the fused FP64 path), written by ESTAB
// in place of its element-order outputs: per element one record of nu =
// SHOC*CTE and tau1..3, and fac = area*dtl/3; element e's values sit at
// patch position inv_perm[e].
template <class R>
struct ElemP {
    R* nt = nullptr;   // 4 per element: nu, tau1, tau2, tau3 (16- or 32-byte records)
    R* fac = nullptr;
    const double *area = nullptr, *dtl = nullptr;  // dtl null: DTL = DTMIN everywhere
    const int* inv_perm = nullptr;  // element e's patch position
    double cte = 1.0;
    // batches: element e belongs to case ecase[e]; its DTMIN is
    // dtmin[ecase[e] * dt_stride] and its constants cases[ecase[e]]
    const int* ecase = nullptr;
    const struct Consts* cases = nullptr;
    int dt_stride = 0;
};
using Elem32 = ElemP<float>;
using Elem64 = ElemP<double>;

// ESTAB (subrutinas.f90:332-446); dtmin is a device scalar.  Element-order
// FP64 outputs SHOC, T_SUGN1..3.
void launch_estab(const Tables& t, const Consts& c, const double* U, const double* T,
                  const double* GAMM, const double* VX, const double* VY, const double* WX,
                  const double* WY, const double* dNx, const double* dNy, const double* dtmin,
                  double* shoc, double* t1, double* t2, double* t3, hipStream_t s);
// The same arithmetic in patch order, writing the patch-order inputs o of
// the fused kernel: one block per patch of ps with the patch's node values
// in LDS.  dN6_patch: the patch-order FP64 gradients, 6 per element (Nx1..3
// then Ny1..3), when they exist (the fused FP64 path), else null (read
// dNx/dNy through ps.perm).
struct PatchSet;
template <class R>
void launch_estab_patch(const Tables& t, const PatchSet& ps, const Consts& c, const double* U,
                        const double* T, const double* GAMM, const double* VX, const double* VY,
                        const double* WX, const double* WY, const double* dNx, const double* dNy,
                        const double* dN6_patch, const double* dtmin, const ElemP<R>& o,
                        hipStream_t s);
// SHOC, T_SUGN1..3 (element order, FP64) from the patch-order copies of o
template <class R>
void launch_estab_widen(const ElemP<R>& o, const int* inv_perm, int nelem, double* shoc,
                        double* t1, double* t2, double* t3, hipStream_t s);

// Element residual: calcRHS (calcRHS.f90:36-151) and, when WX != nullptr,
// FUENTE (subrutinas.f90:1060-1091), written to per-slot buffers indexed
// like esup1 (contribution of element e to node inpoel(k,e) at epos(k,e)).
struct ElemIn {
    const double *U, *theta, *T, *dNx, *dNy, *area, *shoc, *dtl, *t1, *t2, *t3, *WX, *WY;
};
void launch_element_rhs(const Tables& t, const Consts& c, const ElemIn& in, double* c_calc,
                        double* c_fuente, hipStream_t s);

// FUENTE alone into its slot buffer (uses U, dNx, dNy, area, dtl, WX, WY)
void launch_fuente(const Tables& t, const ElemIn& in, double* c_fuente, hipStream_t s);

// Deterministic node gather: rhs(i) = rhs_in(i) + sum calc contributions in
// esup order, then + sum FUENTE contributions (reference accumulation order).
void launch_gather_rhs(const Tables& t, const double* rhs_in, const double* c_calc,
                       const double* c_fuente, double* rhs_out, hipStream_t s);

// Boundary-condition table: unique nodes with their final actions in the
// reference order FIXVEL -> NORMALVEL -> FIX(rho) -> FIX(T).
struct BcTable {
    int nb = 0;
    DBuf<int> bcidx;     // npoin: entry in table or -1
    DBuf<int> flags;     // nb: bit0 vel, bit1 normal, bit2 rho, bit3 T
    DBuf<double> vals;   // nb*8: vx, vy, nx, ny, rho, T, 0, 0
};

// One RK stage node update (subrutinas.f90:695-699, 707-718, 809-826) fused
// with the gather.  final_stage: store U1, RHS and every primitive; otherwise
// only T, the one stage output the next stage reads (every stage evaluates
// the residual of U, subrutinas.f90:676,685, and later stages overwrite the
// rest).
struct StageNodes {
    const double* U;
    double *U1, *RHS, *T, *VX, *VY, *RHO, *E, *P, *RMACH;
    const double *M, *GAMM, *WX, *WY;
};
void launch_stage_update(const Tables& t, const Consts& c, const BcTable& bc, double rk_fact,
                         const double* c_calc, const double* c_fuente, const StageNodes& nd,
                         bool final_stage, hipStream_t s);

// ---------------------------------------------------------------------------
// Element patches for the fused paths (patches.hip): Morton-ordered runs of
// pe elements; per element (in patch order) one 64-bit word (two uint32)
// with the local ids of its nodes (3 x 10 bits) and its local slots (3 x 11
// bits from bit 30); per patch its nodes, interior ones first, with each
// node's slot range (lptr); per halo node its partial-sum entries in
// ascending patch order.
// ---------------------------------------------------------------------------
// default patch shapes, 256 threads each (measured on the 1024 x 2048 O-grid)
constexpr int kPatchElems = 512;     // FP32 path: float4 slots
constexpr int kPatchElems64 = 384;   // FP64 path: double4 slots
constexpr int kPatchMaxNodes = 700;  // keeps the fused kernel's LDS within 64 KB
struct PatchSet {
    int npatch = 0, max_nodes = 0, nentries = 0, nhalo = 0;
    int pe = kPatchElems, block = 256;  // elements per patch, threads per patch
    bool persistent = false;            // a resident grid loops over the patches
    int grid[2] = {0, 0};               // its size, intermediate and final stages
    // batches: case of each patch and of each halo node, the cases' constants
    DBuf<int> pcase, hcase;
    const struct Consts* cases = nullptr;
    DBuf<int> perm, inv_perm;      // patch position <-> element (0-based)
    DBuf<uint32_t> loc;            // 2 per element (patch order)
    DBuf<int> elem_base;           // npatch+1: patch p holds positions [elem_base[p], [p+1])
    DBuf<int> node_base, n_int, halo_base;
    DBuf<int> pnode;               // 1-based node ids, per patch
    DBuf<uint16_t> lptr;           // nl+1 per patch, at node_base[p] + p
    DBuf<int> hnode, hptr, hidx;   // halo nodes (1-based), CSR of their entries
    DBuf<double> hpart;            // 4 per entry
};
void build_patches(PatchSet& ps, const int* inpoel_host, const double* X_host,
                   const double* Y_host, int nelem, int npoin, int pe, int block, hipStream_t s);
// The patch-order numbering of a mesh (host): elem_old[q] = the element
// (0-based) at patch position q; node_old[i] = the node at new position i,
// each patch's nodes in turn (interior nodes, then the halo nodes no earlier
// patch holds).  A mesh renumbered by them gives patches with the same
// elements in the same order, whose interior nodes are contiguous.
void patch_numbering(const int* inpoel, const double* X, const double* Y, int ne, int np, int pe,
                     std::vector<int>& elem_old, std::vector<int>& node_old);
// A batch of disjoint meshes numbered one after the other: each case's
// patches as build_patches makes them, concatenated
void build_patches_batch(PatchSet& ps, int ncases, const int* const* inpoel_host,
                         const double* const* X_host, const double* const* Y_host,
                         const int* nelem, const int* npoin, int pe, int block, hipStream_t s);

// Fused element residual + RK stage for the patch-interior nodes, then the
// stage for the halo nodes (subrutinas.f90:679-826 with calcRHS.f90:37-141
// and FUENTE; element arithmetic in R, node sums and the update in FP64).
// Element inputs are in patch order: dNx/dNy (3 per element), fac, nu,
// tau1..3.
template <class R>
struct PatchIn {
    const double *U, *T, *WX, *WY;
    // per element 4 values (Nx2, Nx3, Ny2, Ny3: the gradients of the second
    // and third shape functions; the first is minus their sum), 16- or
    // 32-byte aligned
    const R* dN;
    const R* fac;
    const R* nt;  // per element nu, tau1, tau2, tau3 (one 16- or 32-byte record)
};
template <class R>
void launch_patch_stage(const Tables& t, const PatchSet& ps, const Consts& c, const BcTable& bc,
                        double rk_fact, const PatchIn<R>& in, const StageNodes& nd,
                        bool final_stage, hipStream_t s);
// A resident grid for ps (persistent = true): as many blocks of the fused
// kernel (element type R) as stay resident on the device at once
template <class R>
void patch_make_persistent(PatchSet& ps, bool fuente);
// whether the fused kernel's LDS (node values, W, slots of slot_bytes,
// slot ranges) fits its 64 KB budget for patches of pe elements and up to
// max_nodes nodes
bool patch_lds_fits(int max_nodes, int pe, size_t slot_bytes);
// dNx/dNy in patch order for the fused FP64 path, from one gather: dN6 holds
// 6 per element (Nx1..3 then Ny1..3, ESTAB's inputs), dN4 the stage kernel's
// 4 (Nx2, Nx3, Ny2, Ny3); both exact
void launch_to_patch64(const double* dNx, const double* dNy, double* dN6, double* dN4,
                       const int* perm, int nelem, hipStream_t s);
// The fused kernel's gradients in patch order, rounded to R: 4 per element
// (Nx2, Nx3, Ny2, Ny3 of the element at each patch position)
template <class R>
void launch_to_patch_grad(const double* dNx, const double* dNy, R* dN, const int* perm, int nelem,
                          hipStream_t s);

// ---------------------------------------------------------------------------
// MLP on the matrix cores (mlp.hip): 16 -> 64 -> 64 -> 16, tanh, tanh,
// linear; kMlpOutUsed outputs are trained.  Parameters (FP32 master copy):
// W1t[64][16], b1[64], W2t[64][64], b2[64], W3t[16][64], b3[16], weights
// stored [out][in].
// ---------------------------------------------------------------------------
constexpr int kMlpIn = 16, kMlpHid = 64, kMlpOut = 16, kMlpOutUsed = 2;
constexpr int kMlpW1 = 0, kMlpB1 = kMlpW1 + kMlpHid * kMlpIn, kMlpW2 = kMlpB1 + kMlpHid,
              kMlpB2 = kMlpW2 + kMlpHid * kMlpHid, kMlpW3 = kMlpB2 + kMlpHid,
              kMlpB3 = kMlpW3 + kMlpOut * kMlpHid, kMlpParams = kMlpB3 + kMlpOut;
struct MlpWeights {
    DBuf<float> w, m, v, grad, loss, gpart, lpart;  // master, Adam moments, gradient
    DBuf<unsigned short> wb;                        // BF16 copy of the weights
    int step = 0;
    void init(const float* host_params, hipStream_t s);
};
// out[2*node + o] (FP32) for the BF16 features X[node][16]
void mlp_forward(const MlpWeights& w, const unsigned short* X, int n, float* out, hipStream_t s);
// gradient of sum_q sum_{o<2} (y - T)^2 / ns over the ns = ceil(n/stride)
// sampled nodes q * stride into w.grad (the undivided sum into w.loss);
// returns ns
int mlp_gradient(MlpWeights& w, const unsigned short* X, const float* T, int n,
                 int tiles_per_block, hipStream_t s, int stride = 1);
void mlp_adam(MlpWeights& w, float lr, hipStream_t s);

// Learned warm start of the mesh solves (warmstart.hip).  WsSolve: the two
// solves' systems and where the session keeps its scalars.
struct PcgWork;
struct WsSolve {
    double *XPOS, *YPOS;            // guesses in, solutions out
    const double *X, *Y;            // node coordinates
    double xref, yref;              // pivot of the body's rotation
    const double* dtmin;            // device scalar
    PcgWork *wx, *wy;
    const double *A, *diag, *b, *fixx, *fixy;
    const int *idx, *rowptr, *fix_idx;
    int nfix;
    double *n_learned, *n_solves, *loss;  // device scalars
};
struct WarmStart {
    int np = 0, nhist = 0;
    bool predicted = false;
    bool use_mlp = true;  // false: the extrapolation alone (the network's ablation)
    bool net_used = false;  // this step's candidate includes the network's correction
    double inv_l = 1.0;
    MlpWeights mlp;
    DBuf<double> h1x, h2x, h3x, h1y, h2y, h3y, ex, ey, hatx, haty, wsc, partials;
    DBuf<unsigned short> feat;
    DBuf<float> out, tgt;
    void init(int npoin, double length, hipStream_t s);
    void before_solves(const WsSolve& q, hipStream_t s);  // candidates and the choice
    void after_solves(const WsSolve& q, hipStream_t s);   // training and history
};

// NORMALVEL (subrutinas.f90:66-85) on the BC table entries NORMALES marked:
// VX, VY in place from W
void launch_normalvel(const BcTable& bc, double* VX, double* VY, const double* WX,
                      const double* WY, hipStream_t s);
// NORMALES (subrutinas.f90:7-64) on the BC table entries that are wall nodes
void launch_normales(const BcTable& bc, int nwallnodes, const int* wnode_bc, const int* wnode_ptr,
                     const int* wnode_edges, const int* wall, const double* X, const double* Y,
                     hipStream_t s);

// Start-up mesh smoothing (smooth.hip, smoothing.f90): the colouring and
// work space of one mesh for a given set of fixed nodes
constexpr int kSmoothSeqMaxNodes = 65536;  // order 0 (one lane) above this: refused
struct SmoothPlan {
    int max_elems = 0;              // most elements around one node
    double hmin = 1.0;              // H_MIN_GLOBAL of the last check
    std::vector<int> colour_ptr;    // free nodes of colour c: colour_nodes[ptr[c], ptr[c+1])
    DBuf<int> colour_nodes, fixed, smoothable;
    DBuf<double> dmove, partials, scal;
    void build(const Tables& t, const int* fixed_host, hipStream_t s);
};
// checkMesh: the smoothable marks and H_MIN_GLOBAL (*hmin_host); false when
// no element is below the quality bound (the reference then returns)
bool smooth_check(const Tables& t, SmoothPlan& plan, double* X, double* Y, double* hmin_host,
                  hipStream_t s);
// the Laplacian passes and the sweeps after smooth_check: X, Y (device) in
// place; tol_move (device, may be null); order 0 = the reference's node
// order, 1 = colour by colour.  Returns the optimisation sweeps run.
int smooth_run(const Tables& t, SmoothPlan& plan, double* X, double* Y, const double* tol_move,
               int order, hipStream_t s);

// ---------------------------------------------------------------------------
// Session state helpers (state.hip), shared by the session and the batch
// ---------------------------------------------------------------------------
// Session scalars live on the device: the CFD_S_* slots, then work slots.
enum { X_RAW = CFD_NSCALARS, X_FDTL, X_CA, X_SA, X_NCTL };

// Boundary-condition table (host construction, reference list semantics)
struct BcHost {
    std::vector<int> nodes, bcidx, flags;
    std::vector<double> vals;
    std::vector<int> wnode_bc, wnode_ptr, wnode_edges;  // wall-node incidence
};
BcHost build_bc_host(int npoin, const cfd_bc_lists* bc);

struct BcDevice {
    BcTable tab;
    DBuf<int> wnode_bc, wnode_ptr, wnode_edges, wall;
    int nw = 0;
    void upload(const BcHost& h, int npoin, const cfd_bc_lists* bc, hipStream_t s) {
        tab.nb = (int)h.nodes.size();
        tab.bcidx.alloc(npoin);
        tab.bcidx.upload(h.bcidx.data(), npoin, s);
        tab.flags.alloc(std::max<size_t>(1, h.flags.size()));
        tab.flags.upload(h.flags.data(), h.flags.size(), s);
        tab.vals.alloc(std::max<size_t>(8, h.vals.size()));
        tab.vals.upload(h.vals.data(), h.vals.size(), s);
        nw = (int)h.wnode_bc.size();
        wnode_bc.alloc(std::max(1, nw));
        wnode_bc.upload(h.wnode_bc.data(), nw, s);
        wnode_ptr.alloc(nw + 1);
        wnode_ptr.upload(h.wnode_ptr.data(), nw + 1, s);
        wnode_edges.alloc(std::max<size_t>(1, h.wnode_edges.size()));
        wnode_edges.upload(h.wnode_edges.data(), h.wnode_edges.size(), s);
        int nwall = bc ? bc->nwall : 0;
        wall.alloc(std::max(2, 2 * nwall));
        if (nwall) wall.upload(bc->wall, 2 * nwall, s);
        CFD_HIP(hipStreamSynchronize(s));
    }
    void normales(const double* X, const double* Y, hipStream_t s) {
        launch_normales(tab, nw, wnode_bc.p, wnode_ptr.p, wnode_edges.p, wall.p, X, Y, s);
    }
};

// free stream (ns2DComp.ALE.f90:409-420) plus the start values of the work
// arrays, for n nodes from the given pointers
struct NodeState {
    double *U, *U1, *VX, *VY, *T, *GAMM, *RHO, *E, *WX, *WY;
};
void launch_freestream(const NodeState& ns, const Consts& c, int n, hipStream_t s);
// the C-ABI constants as the kernels take them
Consts consts_of(const cfd_consts* c);
// Step control: one block per case (ncases scalar blocks of X_NCTL in ctl);
// case c combines partials [pb[c], pb[c+1]) (all nb when pb is null)
void launch_step_ctl(const double* partials, const int* pb, int nb, double* ctl, int ncases,
                     int itlocal, int motion, double amp, double omega, hipStream_t s);
// DT clip and DTL; element e reads case ecase[e] (case 0 when null)
void launch_dt_finish(double* DT, double* DTL, const double* ctl, const int* ecase, bool blend,
                      int nelem, hipStream_t s);

// meshMove.f90 FORCES (out3 = FX, FY, RM; skipped when null) and, when visc
// is given, FORCE_VISC (ns2DComp.ALE.f90:819-893; out2 = F_VX, F_VY) in the
// same launch; device scalars out
struct ViscIn {
    const int* inpoel;
    const double *T, *VX, *VY, *dNx, *dNy;
    double t_inf;
};
// part: device scratch of forces_part_size(nedge) doubles; zero4 (may be
// null): four device scalars set to 0 in the same launches
size_t forces_part_size(int nedge);
void launch_forces(int nedge, const int* iset, const double* X, const double* Y, const double* P,
                   double xref, double yref, double* out3, const ViscIn* visc, double* out2,
                   double* part, double* zero4, hipStream_t s);
// cs: device scalars cos(alpha), sin(alpha)
void launch_transf(int nedge, const int* iset, const double* X, const double* Y, const double* cs,
                   double yposr, double xref, double yref, double* DXPOS, double* DYPOS,
                   hipStream_t s);

// ---------------------------------------------------------------------------
// PCG (biconjGrad.f90:8-62)
// ---------------------------------------------------------------------------
struct PcgWork {
    int n = 0, nblocks = 0;
    DBuf<double> r, z, y, pa, pb, partial;
    DBuf<double> scal;         // see pcg.hip for the slot meaning
    DBuf<int> fixflag;         // n: 1 on Dirichlet nodes
    HostPinned<double> hpin;   // host copy of scal for the polls (null: a stack copy)
    void ensure(int npoin);
};
// One system of a multi-system solve: its own work space, x, b and fixed
// values; the matrix is shared.
struct PcgSystem {
    PcgWork* w;
    double* x;
    const double* b;
    const double* x_fix;
    const int* x_fix_idx;
    int nfix;
};
// The matrix of a CSR system stored a second time in ELL order for the
// SpMV kernels: entry k of row i (its k-th CSR entry) at k * n + i, column
// 0-based; rows keep their CSR lengths (entries beyond are never read), so a
// row's products are summed in the same order as from the CSR arrays.
constexpr int kEllMaxWidth = 16;
struct EllMat {
    const double* val = nullptr;
    const int* col = nullptr;
    const int* rowptr = nullptr;
    int width = 0;
};
struct EllLap {
    int n = 0, width = 0;
    DBuf<int> col;
    DBuf<double> val;
    // builds the column layout; false (nothing built) when the longest row
    // exceeds kEllMaxWidth or 1.5 times the mean row length
    bool build(const Tables& t, hipStream_t s);
    // val from the CSR values A
    void refresh(const Tables& t, const double* A, hipStream_t s);
    EllMat mat(const Tables& t) const { return EllMat{val.p, col.p, t.rowptr.p, width}; }
};
// Several independent solves on one matrix, interleaved on stream s with
// shared polls; ell (when given) is the same matrix in ELL order, used by
// the SpMV kernels of the iterations (bit for bit the CSR kernels' results).
void pcg_launch_multi(const PcgSystem* sys, int nsys, int npoin, const double* A, const int* idx,
                      const int* rowptr, const double* diag, double tol, int maxit, int rel,
                      hipStream_t s, const EllMat* ell = nullptr);
// iterations of the last solve run by pcg_launch_multi on w, as its final
// poll read them (host memory, no device access)
double pcg_polled_iterations(const PcgWork& w);
struct PcgResult {
    int iters;
    double rz;
    int converged;
};
// One solve on stream s; the host polls the done flag between batches of
// iterations, so the call returns once the loop has ended on the device.
// x (device) in/out; b (device); x_fix/x_fix_idx device arrays of nfix.
// rel = 0: the reference's absolute tests against tol; rel = 1: relative
// (see pcg.hip).
void pcg_launch(PcgWork& w, int npoin, const double* A, const int* idx, const int* rowptr,
                const double* diag, double* x, const double* b, const double* x_fix,
                const int* x_fix_idx, int nfix, double tol, int maxit, int rel, hipStream_t s);
// Waits for the stream and returns the last solve's outcome.
PcgResult pcg_result(const PcgWork& w, hipStream_t s);
// Copy the last solve's iteration count and final r.z to device scalars
// (and add the count to *itsum when given, and 1 to *unconv when given and
// the solve stopped at its iteration cap without meeting the tolerance).
void pcg_store_result(const PcgWork& w, double* iters, double* rz, double* itsum, hipStream_t s,
                      double* unconv = nullptr);
// r.r of the first residual test for the guess x (fixed values applied to a
// copy; x is not changed) into the device scalar rr_out
void pcg_initial_rr(PcgWork& w, int npoin, const double* A, const int* idx, const int* rowptr,
                    const double* diag, const double* x, const double* b, const double* x_fix,
                    const int* x_fix_idx, int nfix, double* rr_out, hipStream_t s);

}  // namespace cfd
