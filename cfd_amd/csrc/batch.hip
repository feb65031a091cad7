// Batches of independent fixed-mesh cases of uneven size in one device state.
//
// The cases are numbered one after the other into one disjoint union mesh:
// case k owns nodes [nbase[k], nbase[k+1]) and elements [ebase[k],
// ebase[k+1]).  Each case keeps its own step control (DTMIN, its 1 %
// hysteresis, the clock: X_NCTL scalars per case) and its own free stream
// (rho_inf, u_inf, v_inf: the Mach number, angle of attack and Reynolds
// number of a sweep); every other constant is shared.  Per RK stage one
// launch of the fused patch kernel covers every case (a resident grid loops
// over all cases' patches, whose shape -- elements and threads per patch --
// is a parameter the autotuner picks), followed by the halo pass.  Patches
// never straddle two cases and are built per case exactly as a single-case
// session builds them, so each case's arithmetic and summation order are
// those of that case run on its own: results agree bit for bit.
//
// A whole time step (deltat with one min partial per case chunk, step
// control per case, ESTAB, the four fused stages) is recorded once into a
// hipGraph per U buffer parity and replayed; nothing in it synchronises
// with the host or clears memory with hipMemsetAsync.
#include <map>
#include <memory>

#include "cfd_internal.hpp"

using namespace cfd;

namespace {
constexpr int kChunkElems = 8192;  // elements per deltat partial
}

struct cfd_batch {
    hipStream_t s = nullptr;
    int ncases = 0, precision = 1, itlocal = 0, nrk = 4;
    int np = 0, ne = 0, nchunks = 0;
    std::vector<int> nbase, ebase;
    std::vector<Consts> cc;  // per case
    Consts c{};              // the shared constants (case 0's free stream)
    Tables t;
    BcDevice bc;
    PatchSet patches;
    DBuf<double> Ub[2];
    int ucur = 0;
    DBuf<double> RHS, T, VX, VY, WX, WY, X, Y, RHO, E, P, RMACH, M, GAMM;
    DBuf<double> DTL, DT, AREA, DNX, DNY, HH, HHX, HHY, SHOC, T1, T2, T3;
    DBuf<float> DN32, NT32, FAC32;
    DBuf<double> DN6, DN64, NT64, FAC64;
    DBuf<int> ecase, chunk, chunk_case, pb;
    DBuf<Consts> cases;  // per case
    DBuf<double> ctl, partials, hmin;
    // host copies for re-building patches (autotuning)
    std::vector<std::vector<int>> inpoel_h;   // per case, in patch-order numbering
    std::vector<std::vector<double>> X_h, Y_h;
    // per case: the caller's node / element at each patch-order position
    std::vector<std::vector<int>> node_old, elem_old;
    // graphs: one recorded step per U buffer parity
    bool use_graph = true;
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    hipGraph_t graph[2] = {nullptr, nullptr};

    double* U() { return Ub[ucur].p; }
    double* U1() { return Ub[ucur ^ 1].p; }

    ~cfd_batch() {
        if (s) (void)hipStreamSynchronize(s);
        drop_graphs();
        if (s) (void)hipStreamDestroy(s);
    }
    void drop_graphs() {
        for (int k = 0; k < 2; ++k) {
            if (gexec[k]) (void)hipGraphExecDestroy(gexec[k]);
            if (graph[k]) (void)hipGraphDestroy(graph[k]);
            gexec[k] = nullptr;
            graph[k] = nullptr;
        }
    }

    template <class R>
    ElemP<R> elem_patch(R* nt, R* fac) {
        ElemP<R> o;
        o.nt = nt;
        o.fac = fac;
        o.area = AREA.p;
        o.dtl = DTL.p;
        o.inv_perm = patches.inv_perm.p;
        o.cte = c.cte;
        o.ecase = ecase.p;
        o.cases = cases.p;
        o.dt_stride = X_NCTL;
        return o;
    }
    Elem32 elem32() { return elem_patch(NT32.p, FAC32.p); }
    Elem64 elem64() { return elem_patch(NT64.p, FAC64.p); }

    void refresh_patch_geometry() {
        if (precision == 1) {
            launch_to_patch_grad(DNX.p, DNY.p, DN32.p, patches.perm.p, ne, s);
        } else {
            launch_to_patch64(DNX.p, DNY.p, DN6.p, DN64.p, patches.perm.p, ne, s);
        }
    }

    void build(int pe, int block) {
        std::vector<const int*> ip(ncases);
        std::vector<const double*> xs(ncases), ys(ncases);
        std::vector<int> nes(ncases), nps(ncases);
        for (int k = 0; k < ncases; ++k) {
            ip[k] = inpoel_h[k].data();
            xs[k] = X_h[k].data();
            ys[k] = Y_h[k].data();
            nes[k] = ebase[k + 1] - ebase[k];
            nps[k] = nbase[k + 1] - nbase[k];
        }
        build_patches_batch(patches, ncases, ip.data(), xs.data(), ys.data(), nes.data(), nps.data(),
                            pe, block, s);
        patches.cases = cases.p;
        if (precision == 1)
            patch_make_persistent<float>(patches, false);
        else
            patch_make_persistent<double>(patches, false);
        refresh_patch_geometry();
        drop_graphs();
    }

    // One time step of every case (ns2DComp.ALE.f90:140-182 on a fixed mesh)
    void step_body() {
        const double* dtmin = ctl.p + CFD_S_DTMIN;
        launch_deltat_chunks(t, c, AREA.p, T.p, VX.p, VY.p, WX.p, WY.p, DT.p, chunk.p, nchunks,
                             partials.p, chunk_case.p, cases.p, s);
        launch_step_ctl(partials.p, pb.p, nchunks, ctl.p, ncases, itlocal, 0, 0.0, 0.0, s);
        launch_dt_finish(DT.p, DTL.p, ctl.p, ecase.p, itlocal != 0, ne, s);
        if (precision == 1)
            launch_estab_patch(t, patches, c, U(), T.p, GAMM.p, VX.p, VY.p, WX.p, WY.p, DNX.p,
                               DNY.p, nullptr, dtmin, elem32(), s);
        else
            launch_estab_patch(t, patches, c, U(), T.p, GAMM.p, VX.p, VY.p, WX.p, WY.p, DNX.p,
                               DNY.p, DN6.p, dtmin, elem64(), s);
        PatchIn<float> in32{U(), T.p, nullptr, nullptr, DN32.p,
                            FAC32.p, NT32.p};
        PatchIn<double> in64{U(), T.p, nullptr, nullptr, DN64.p,
                             FAC64.p, NT64.p};
        StageNodes sn{U(), U1(), RHS.p, T.p, VX.p, VY.p, RHO.p, E.p, P.p, RMACH.p, M.p, GAMM.p,
                      WX.p, WY.p};
        for (int irk = 1; irk <= nrk; ++irk) {
            const double rk_fact = 1.0 / (nrk + 1 - irk);
            if (precision == 1)
                launch_patch_stage(t, patches, c, bc.tab, rk_fact, in32, sn, irk == nrk, s);
            else
                launch_patch_stage(t, patches, c, bc.tab, rk_fact, in64, sn, irk == nrk, s);
        }
    }

    void step(int n) {
        for (int k = 0; k < n; ++k) {
            if (!use_graph) {
                step_body();
            } else {
                if (!gexec[ucur]) {
                    CFD_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                    try {
                        step_body();
                    } catch (...) {
                        hipGraph_t g = nullptr;
                        (void)hipStreamEndCapture(s, &g);
                        if (g) (void)hipGraphDestroy(g);
                        throw;
                    }
                    CFD_HIP(hipStreamEndCapture(s, &graph[ucur]));
                    CFD_HIP(hipGraphInstantiate(&gexec[ucur], graph[ucur], nullptr, nullptr, 0));
                }
                CFD_HIP(hipGraphLaunch(gexec[ucur], s));
            }
            ucur ^= 1;  // U = U1
        }
    }

    void init_freestream() {
        for (int k = 0; k < ncases; ++k) {
            const size_t o = nbase[k];
            launch_freestream({U() + 4 * o, U1() + 4 * o, VX.p + o, VY.p + o, T.p + o, GAMM.p + o,
                               RHO.p + o, E.p + o, WX.p + o, WY.p + o},
                              cc[k], nbase[k + 1] - nbase[k], s);
        }
        std::vector<double> sc((size_t)ncases * X_NCTL, 0.0);
        for (int k = 0; k < ncases; ++k) sc[(size_t)k * X_NCTL + CFD_S_BANDERA] = 1.0;
        ctl.upload(sc.data(), sc.size(), s);
        CFD_HIP(hipStreamSynchronize(s));
    }

    // (buffer, values per node or element, node field?)
    DBuf<double>* field(int f, int& width, bool& node) {
        node = true;
        width = 1;
        switch (f) {
            case CFD_F_U: width = 4; return &Ub[ucur];
            case CFD_F_U1: width = 4; return &Ub[ucur ^ 1];
            case CFD_F_RHS: width = 4; return &RHS;
            case CFD_F_T: return &T;
            case CFD_F_VEL_X: return &VX;
            case CFD_F_VEL_Y: return &VY;
            case CFD_F_W_X: return &WX;
            case CFD_F_W_Y: return &WY;
            case CFD_F_X: return &X;
            case CFD_F_Y: return &Y;
            case CFD_F_RHO: return &RHO;
            case CFD_F_E: return &E;
            case CFD_F_P: return &P;
            case CFD_F_RMACH: return &RMACH;
            case CFD_F_M: return &M;
            case CFD_F_GAMM: return &GAMM;
            default: break;
        }
        node = false;
        switch (f) {
            case CFD_F_DTL: return &DTL;
            case CFD_F_DT: return &DT;
            case CFD_F_SHOC: return &SHOC;
            case CFD_F_T_SUGN1: return &T1;
            case CFD_F_T_SUGN2: return &T2;
            case CFD_F_T_SUGN3: return &T3;
            case CFD_F_AREA: return &AREA;
            case CFD_F_DNX: width = 3; return &DNX;
            case CFD_F_DNY: width = 3; return &DNY;
            case CFD_F_HH: return &HH;
            case CFD_F_HHX: return &HHX;
            case CFD_F_HHY: return &HHY;
            default: throw Error("arg: field not kept by a batch");
        }
    }
};




extern "C" {

int cfd_batch_create(cfd_batch** out, int ncases, const cfd_batch_case* cases, int precision,
                     int itlocal, int patch_elems, int patch_block) {
    CFD_TRY({
        CFD_ARG(out && cases && ncases > 0, "batch arguments");
        CFD_ARG(precision == 0 || precision == 1, "batch precision must be 0 (FP64) or 1 (FP32)");
        int ndev = 0;
        CFD_HIP(hipGetDeviceCount(&ndev));
        CFD_CHECK(ndev > 0, "no GPU");
        auto B = std::make_unique<cfd_batch>();
        CFD_HIP(hipStreamCreate(&B->s));
        hipStream_t s = B->s;
        B->ncases = ncases;
        B->precision = precision;
        B->itlocal = itlocal;
        B->nbase.assign(1, 0);
        B->ebase.assign(1, 0);
        long long np = 0, ne = 0;
        for (int k = 0; k < ncases; ++k) {
            const cfd_batch_case& q = cases[k];
            CFD_ARG(q.nelem > 0 && q.npoin > 0 && q.inpoel && q.X && q.Y, "case mesh");
            np += q.npoin;
            ne += q.nelem;
            CFD_ARG(np < (1ll << 31) / 4 && ne < (1ll << 31) / 3, "batch too large");
            B->nbase.push_back((int)np);
            B->ebase.push_back((int)ne);
            B->cc.push_back(consts_of(&q.c));
        }
        B->np = (int)np;
        B->ne = (int)ne;
        B->c = B->cc[0];
        const int pe = patch_elems > 0 ? patch_elems : (precision == 1 ? kPatchElems : kPatchElems64);
        const int blk = patch_block > 0 ? patch_block : 256;
        CFD_ARG(pe >= 1 && pe <= 682, "patch_elems must be 1..682");
        CFD_ARG(blk % 64 == 0 && blk >= 64 && blk <= 512,
                "patch_block must be a multiple of 64 from 64 to 512");
        // each case numbered in patch order inside (elements in patch order,
        // each patch's nodes in turn): the same results bit for bit as the
        // caller's numbering on these fixed meshes, with contiguous patch nodes
        B->node_old.resize(ncases);
        B->elem_old.resize(ncases);
        std::vector<std::vector<int>> node_new(ncases);
        for (int k = 0; k < ncases; ++k) {
            const cfd_batch_case& q = cases[k];
            for (long long i = 0; i < 3ll * q.nelem; ++i)
                CFD_CHECK(q.inpoel[i] >= 1 && q.inpoel[i] <= q.npoin, "mesh: node id out of range");
            patch_numbering(q.inpoel, q.X, q.Y, q.nelem, q.npoin, pe, B->elem_old[k], B->node_old[k]);
            node_new[k].assign(q.npoin, 0);
            for (int i = 0; i < q.npoin; ++i) node_new[k][B->node_old[k][i]] = i;
        }
        // the union mesh, coordinates and boundary lists (ids shifted per case)
        std::vector<int> inp(3 * (size_t)ne);
        std::vector<double> X(np), Y(np);
        std::vector<int> fv, fr, ft, wall;
        std::vector<double> fvx, fvy, frv, ftv;
        B->inpoel_h.resize(ncases);
        B->X_h.resize(ncases);
        B->Y_h.resize(ncases);
        for (int k = 0; k < ncases; ++k) {
            const cfd_batch_case& q = cases[k];
            const int N = B->nbase[k], E = B->ebase[k];
            const std::vector<int>& nn = node_new[k];
            const std::vector<int>& eo = B->elem_old[k];
            std::vector<int>& lin = B->inpoel_h[k];
            lin.resize(3 * (size_t)q.nelem);
            for (int e = 0; e < q.nelem; ++e)
                for (int j = 0; j < 3; ++j) {  // each element keeps its node order
                    const int n = nn[q.inpoel[3 * (size_t)eo[e] + j] - 1] + 1;
                    lin[3 * (size_t)e + j] = n;
                    inp[3 * ((size_t)E + e) + j] = n + N;
                }
            B->X_h[k].resize(q.npoin);
            B->Y_h[k].resize(q.npoin);
            for (int i = 0; i < q.npoin; ++i) {
                B->X_h[k][i] = X[N + i] = q.X[B->node_old[k][i]];
                B->Y_h[k][i] = Y[N + i] = q.Y[B->node_old[k][i]];
            }
            const cfd_bc_lists& b = q.bc;
            auto shift = [&](int n) {
                CFD_CHECK(n >= 1 && n <= q.npoin, "boundary list node id out of range");
                return nn[n - 1] + 1 + N;
            };
            for (int i = 0; i < b.nfixv; ++i) {
                fv.push_back(shift(b.ifixv_node[i]));
                fvx.push_back(b.rfixv_valuex[i]);
                fvy.push_back(b.rfixv_valuey[i]);
            }
            for (int i = 0; i < b.nfixrho; ++i) {
                fr.push_back(shift(b.ifixrho_node[i]));
                frv.push_back(b.rfixrho_value[i]);
            }
            for (int i = 0; i < b.nfixt; ++i) {
                ft.push_back(shift(b.ifixt_node[i]));
                ftv.push_back(b.rfixt_value[i]);
            }
            for (int i = 0; i < 2 * b.nwall; ++i) wall.push_back(shift(b.wall[i]));
        }
        cfd_bc_lists ub{};
        ub.nfixv = (int)fv.size();
        ub.ifixv_node = fv.data();
        ub.rfixv_valuex = fvx.data();
        ub.rfixv_valuey = fvy.data();
        ub.nfixrho = (int)fr.size();
        ub.ifixrho_node = fr.data();
        ub.rfixrho_value = frv.data();
        ub.nfixt = (int)ft.size();
        ub.ifixt_node = ft.data();
        ub.rfixt_value = ftv.data();
        ub.nwall = (int)wall.size() / 2;
        ub.wall = wall.data();
        build_tables(B->t, inp.data(), B->ne, B->np, s);
        {  // per-node sums (MASAS) in each case's original element order
            std::vector<int> key(ne);
            for (int k = 0; k < ncases; ++k)
                for (int e = 0; e < cases[k].nelem; ++e)
                    key[B->ebase[k] + e] = B->ebase[k] + B->elem_old[k][e];
            reorder_esup(B->t, key, s);
        }
        B->bc.upload(build_bc_host(B->np, &ub), B->np, &ub, s);
        // state
        for (auto* b : {&B->Ub[0], &B->Ub[1], &B->RHS}) b->alloc(4 * (size_t)np);
        for (auto* b : {&B->T, &B->VX, &B->VY, &B->WX, &B->WY, &B->X, &B->Y, &B->RHO, &B->E, &B->P,
                        &B->RMACH, &B->M, &B->GAMM})
            b->alloc(np);
        for (auto* b : {&B->DTL, &B->DT, &B->AREA, &B->HH, &B->HHX, &B->HHY, &B->SHOC, &B->T1,
                        &B->T2, &B->T3})
            b->alloc(ne);
        B->DNX.alloc(3 * (size_t)ne);
        B->DNY.alloc(3 * (size_t)ne);
        if (precision == 1) {
            B->DN32.alloc(4 * (size_t)ne);
            B->NT32.alloc(4 * (size_t)ne);
            B->FAC32.alloc(ne);
        } else {
            B->DN6.alloc(6 * (size_t)ne);
            B->DN64.alloc(4 * (size_t)ne);
            B->NT64.alloc(4 * (size_t)ne);
            B->FAC64.alloc(ne);
        }
        for (auto* b : {&B->Ub[0], &B->Ub[1], &B->RHS, &B->P, &B->RMACH, &B->DTL, &B->DT})
            b->zero(s);
        B->X.upload(X.data(), np, s);
        B->Y.upload(Y.data(), np, s);
        // per case: element -> case, deltat chunks inside cases, rho_inf
        std::vector<int> ec(ne), ch(1, 0), chc, cpb(1, 0);
        for (int k = 0; k < ncases; ++k) {
            for (int e = B->ebase[k]; e < B->ebase[k + 1]; ++e) ec[e] = k;
            for (int e = B->ebase[k]; e < B->ebase[k + 1]; e += kChunkElems) {
                ch.push_back(std::min(e + kChunkElems, B->ebase[k + 1]));
                chc.push_back(k);
            }
            cpb.push_back((int)ch.size() - 1);
        }
        B->nchunks = (int)ch.size() - 1;
        B->ecase.alloc(ne);
        B->ecase.upload(ec.data(), ne, s);
        B->chunk.alloc(ch.size());
        B->chunk.upload(ch.data(), ch.size(), s);
        B->chunk_case.alloc(chc.size());
        B->chunk_case.upload(chc.data(), chc.size(), s);
        B->pb.alloc(cpb.size());
        B->pb.upload(cpb.data(), cpb.size(), s);
        B->cases.alloc(ncases);
        B->cases.upload(B->cc.data(), ncases, s);
        B->ctl.alloc((size_t)ncases * X_NCTL);
        B->ctl.zero(s);
        B->partials.alloc(std::max<size_t>(grid_for(ne), B->nchunks));
        B->hmin.alloc(1);
        // geometry once (fixed meshes): NORMALES, DERIV, MASAS
        B->bc.normales(B->X.p, B->Y.p, s);
        launch_deriv(B->t, B->X.p, B->Y.p, B->AREA.p, B->DNX.p, B->DNY.p, B->HH.p, B->HHX.p,
                     B->HHY.p, B->hmin.p, B->partials.p, s);
        launch_masas(B->t, B->AREA.p, B->M.p, s);
        B->build(pe, blk);
        CFD_ARG(patch_lds_fits(B->patches.max_nodes, pe,
                               precision == 1 ? sizeof(float4) : sizeof(double4)),
                "patch_elems too large: a patch exceeds the fused kernel's 64 KB of LDS");
        B->init_freestream();
        CFD_HIP(hipStreamSynchronize(s));
        *out = B.release();
    })
}

int cfd_batch_destroy(cfd_batch* b) {
    CFD_TRY({ delete b; })
}

int cfd_batch_init_freestream(cfd_batch* b) {
    CFD_TRY({
        CFD_ARG(b, "null batch");
        b->ucur = 0;
        b->init_freestream();
    })
}

int cfd_batch_set_graph(cfd_batch* b, int use_graph) {
    CFD_TRY({
        CFD_ARG(b, "null batch");
        CFD_HIP(hipStreamSynchronize(b->s));
        b->use_graph = use_graph != 0;
    })
}

int cfd_batch_step(cfd_batch* b, int nsteps) {
    CFD_TRY({
        CFD_ARG(b && nsteps >= 0, "arguments");
        b->step(nsteps);
    })
}

int cfd_batch_sync(cfd_batch* b) {
    CFD_TRY({
        CFD_ARG(b, "null batch");
        CFD_HIP(hipStreamSynchronize(b->s));
    })
}

int cfd_batch_case_size(cfd_batch* b, int icase, int field, long long* count) {
    CFD_TRY({
        CFD_ARG(b && count && icase >= 0 && icase < b->ncases, "arguments");
        int w = 1;
        bool node = true;
        b->field(field, w, node);
        const int n = node ? b->nbase[icase + 1] - b->nbase[icase] : b->ebase[icase + 1] - b->ebase[icase];
        *count = (long long)w * n;
    })
}

int cfd_batch_get_field(cfd_batch* b, int icase, int field, double* host) {
    CFD_TRY({
        CFD_ARG(b && host && icase >= 0 && icase < b->ncases, "arguments");
        int w = 1;
        bool node = true;
        DBuf<double>* buf = b->field(field, w, node);
        const bool estab = field == CFD_F_SHOC || field == CFD_F_T_SUGN1 ||
                           field == CFD_F_T_SUGN2 || field == CFD_F_T_SUGN3;
        if (estab) {  // the values the fused kernel used, back in element order
            if (b->precision == 1)
                launch_estab_widen(b->elem32(), b->patches.inv_perm.p, b->ne, b->SHOC.p, b->T1.p,
                                   b->T2.p, b->T3.p, b->s);
            else
                launch_estab_widen(b->elem64(), b->patches.inv_perm.p, b->ne, b->SHOC.p, b->T1.p,
                                   b->T2.p, b->T3.p, b->s);
        }
        const size_t o = node ? b->nbase[icase] : b->ebase[icase];
        const size_t n = node ? b->nbase[icase + 1] - o : b->ebase[icase + 1] - o;
        std::vector<double> tmp(w * n);
        CFD_HIP(hipMemcpyAsync(tmp.data(), buf->p + w * o, sizeof(double) * w * n,
                               hipMemcpyDeviceToHost, b->s));
        CFD_HIP(hipStreamSynchronize(b->s));
        // back to the caller's numbering of the case
        const std::vector<int>& old = node ? b->node_old[icase] : b->elem_old[icase];
        for (size_t i = 0; i < n; ++i)
            for (int c = 0; c < w; ++c) host[w * (size_t)old[i] + c] = tmp[w * i + c];
    })
}

int cfd_batch_get_scalars(cfd_batch* b, int icase, double* out) {
    CFD_TRY({
        CFD_ARG(b && out && icase >= 0 && icase < b->ncases, "arguments");
        CFD_HIP(hipMemcpyAsync(out, b->ctl.p + (size_t)icase * X_NCTL, sizeof(double) * CFD_NSCALARS,
                               hipMemcpyDeviceToHost, b->s));
        CFD_HIP(hipStreamSynchronize(b->s));
    })
}

int cfd_batch_info(cfd_batch* b, long long* npoin, long long* nelem, int* npatch, int* patch_elems,
                   int* patch_block) {
    CFD_TRY({
        CFD_ARG(b, "null batch");
        if (npoin) *npoin = b->np;
        if (nelem) *nelem = b->ne;
        if (npatch) *npatch = b->patches.npatch;
        if (patch_elems) *patch_elems = b->patches.pe;
        if (patch_block) *patch_block = b->patches.block;
    })
}

int cfd_batch_autotune(cfd_batch* b, int ncand, const int* patch_elems, const int* patch_block,
                       int nsteps, double* ms_per_step, int* best) {
    CFD_TRY({
        CFD_ARG(b && ncand > 0 && patch_elems && patch_block && nsteps > 0, "arguments");
        const size_t slot = b->precision == 1 ? sizeof(float4) : sizeof(double4);
        int ibest = -1;
        double tbest = 0.0;
        hipEvent_t e0, e1;
        CFD_HIP(hipEventCreate(&e0));
        CFD_HIP(hipEventCreate(&e1));
        struct EG { hipEvent_t a, b; ~EG() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); } } eg{e0, e1};
        for (int i = 0; i < ncand; ++i) {
            const int pe = patch_elems[i], blk = patch_block[i];
            CFD_ARG(pe >= 1 && pe <= 682 && blk % 64 == 0 && blk >= 64 && blk <= 512,
                    "candidate patch shape");
            double ms = -1.0;  // -1: the shape does not fit the fused kernel's LDS
            bool fits = patch_lds_fits(0, pe, slot);
            if (fits) {
                try {
                    b->build(pe, blk);
                } catch (const cfd::Error&) {  // a patch with too many nodes
                    fits = false;
                }
                fits = fits && patch_lds_fits(b->patches.max_nodes, pe, slot);
            }
            if (fits) {
                b->ucur = 0;
                b->init_freestream();
                b->step(2);  // records the graphs, warms up
                CFD_HIP(hipEventRecord(e0, b->s));
                b->step(nsteps);
                CFD_HIP(hipEventRecord(e1, b->s));
                CFD_HIP(hipEventSynchronize(e1));
                float f = 0.f;
                CFD_HIP(hipEventElapsedTime(&f, e0, e1));
                ms = f / nsteps;
                if (ibest < 0 || ms < tbest) {
                    ibest = i;
                    tbest = ms;
                }
            }
            if (ms_per_step) ms_per_step[i] = ms;
        }
        CFD_CHECK(ibest >= 0, "arg: no candidate patch shape fits");
        b->build(patch_elems[ibest], patch_block[ibest]);
        b->ucur = 0;
        b->init_freestream();
        if (best) *best = ibest;
    })
}

}  // extern "C"
