// Element patches for the fused residual + stage kernels (FP32 and FP64).
//
// Elements are ordered along a Morton (Z-order) curve of their centroids and
// cut into patches of pe consecutive elements, so a patch is a
// compact piece of the mesh.  A node is interior to a patch when every
// element around it lies in that patch; otherwise it is a halo node, and
// each patch that touches it keeps one partial sum for it.
//
// Per patch, the fused kernel needs: its nodes (interior first, each group in
// ascending global id), and per element the local ids of its three nodes and
// the local slot each contribution goes to.  A node's slots are contiguous
// and in ascending element order inside the patch (lptr), so the node sums
// them in a fixed order; a halo node's partials are summed in ascending patch
// order.  Everything is deterministic.
#include <numeric>

#include "cfd_internal.hpp"

namespace cfd {

namespace {
inline uint64_t spread_bits(uint64_t v) {  // bit i -> bit 2i (32-bit input)
    v &= 0xFFFFFFFFull;
    v = (v | (v << 16)) & 0x0000FFFF0000FFFFull;
    v = (v | (v << 8)) & 0x00FF00FF00FF00FFull;
    v = (v | (v << 4)) & 0x0F0F0F0F0F0F0F0Full;
    v = (v | (v << 2)) & 0x3333333333333333ull;
    v = (v | (v << 1)) & 0x5555555555555555ull;
    return v;
}
}  // namespace

// One mesh's patches, host side.  Node and element ids are the mesh's own.
struct PatchHost {
    int npatch = 0, max_nodes = 0, nentries = 0;
    std::vector<int> perm, elem_base, node_base, n_int, halo_base, pnode, hnode, hptr, hidx;
    std::vector<uint16_t> lptr;
    std::vector<uint32_t> loc;
};

static PatchHost patches_host(const int* inpoel, const double* X, const double* Y, int ne, int np,
                              int pe) {
    CFD_CHECK(ne > 0 && np > 0, "mesh: empty mesh");
    // Morton order of element centroids over the bounding box
    double xmin = X[0], xmax = X[0], ymin = Y[0], ymax = Y[0];
    for (int i = 1; i < np; ++i) {
        xmin = std::min(xmin, X[i]);
        xmax = std::max(xmax, X[i]);
        ymin = std::min(ymin, Y[i]);
        ymax = std::max(ymax, Y[i]);
    }
    const double sx = xmax > xmin ? 2147483647.0 / (xmax - xmin) : 0.0;
    const double sy = ymax > ymin ? 2147483647.0 / (ymax - ymin) : 0.0;
    std::vector<std::pair<uint64_t, int>> key(ne);
    for (int e = 0; e < ne; ++e) {
        const int* ip = inpoel + 3 * (size_t)e;
        double cx = 0.0, cy = 0.0;
        for (int k = 0; k < 3; ++k) {
            const int n = ip[k] - 1;
            CFD_CHECK(n >= 0 && n < np, "mesh: node id out of range");
            cx += X[n];
            cy += Y[n];
        }
        cx /= 3.0;
        cy /= 3.0;
        const uint64_t qx = (uint64_t)((cx - xmin) * sx), qy = (uint64_t)((cy - ymin) * sy);
        key[e] = {spread_bits(qx) | (spread_bits(qy) << 1), e};
    }
    std::sort(key.begin(), key.end());
    std::vector<int> perm(ne);
    for (int q = 0; q < ne; ++q) perm[q] = key[q].second;
    key.clear();
    key.shrink_to_fit();

    const int npatch = (ne + pe - 1) / pe;
    // number of distinct patches around each node
    std::vector<int> npat(np, 0), last(np, -1);
    for (int p = 0; p < npatch; ++p) {
        const int q1 = std::min(ne, (p + 1) * pe);
        for (int q = p * pe; q < q1; ++q)
            for (int k = 0; k < 3; ++k) {
                const int n = inpoel[3 * (size_t)perm[q] + k] - 1;
                if (last[n] != p) {
                    last[n] = p;
                    ++npat[n];
                }
            }
    }
    std::vector<int> node_base(npatch + 1, 0), n_int(npatch), halo_base(npatch + 1, 0);
    std::vector<int> pnode;
    std::vector<uint16_t> lptr;  // nl+1 entries per patch, at node_base[p] + p
    std::vector<uint32_t> loc(2 * (size_t)ne);
    std::vector<int> lid(np, -1), stamp(np, -1), cnt, cursor;
    std::vector<int> hnode_of_entry;  // global node of each halo entry
    pnode.reserve((size_t)np * 5 / 4);
    lptr.reserve((size_t)np * 5 / 4 + npatch);
    int max_nodes = 0;
    for (int p = 0; p < npatch; ++p) {
        const int q0 = p * pe, q1 = std::min(ne, (p + 1) * pe);
        std::vector<int> in_nodes, halo_nodes;
        for (int q = q0; q < q1; ++q)
            for (int k = 0; k < 3; ++k) {
                const int n = inpoel[3 * (size_t)perm[q] + k] - 1;
                if (stamp[n] != p) {
                    stamp[n] = p;
                    (npat[n] == 1 ? in_nodes : halo_nodes).push_back(n);
                }
            }
        std::sort(in_nodes.begin(), in_nodes.end());
        std::sort(halo_nodes.begin(), halo_nodes.end());
        const int nl = (int)(in_nodes.size() + halo_nodes.size());
        CFD_CHECK(nl <= kPatchMaxNodes, "mesh: an element patch has too many nodes");
        max_nodes = std::max(max_nodes, nl);
        n_int[p] = (int)in_nodes.size();
        node_base[p + 1] = node_base[p] + nl;
        halo_base[p + 1] = halo_base[p] + (int)halo_nodes.size();
        int l = 0;
        for (int n : in_nodes) { lid[n] = l++; pnode.push_back(n + 1); }
        for (int n : halo_nodes) { lid[n] = l++; pnode.push_back(n + 1); hnode_of_entry.push_back(n); }
        // slot layout: each node's contributions contiguous, ascending element
        cnt.assign(nl, 0);
        for (int q = q0; q < q1; ++q)
            for (int k = 0; k < 3; ++k) ++cnt[lid[inpoel[3 * (size_t)perm[q] + k] - 1]];
        cursor.assign(nl, 0);
        int acc = 0;
        for (int j = 0; j < nl; ++j) {
            lptr.push_back((uint16_t)acc);
            cursor[j] = acc;
            acc += cnt[j];
        }
        lptr.push_back((uint16_t)acc);
        for (int q = q0; q < q1; ++q) {  // 3 x 10-bit node ids, then 3 x 11-bit slots
            uint64_t w = 0;
            for (int k = 0; k < 3; ++k) {
                const int j = lid[inpoel[3 * (size_t)perm[q] + k] - 1];
                w |= (uint64_t)j << (10 * k);
                w |= (uint64_t)cursor[j]++ << (30 + 11 * k);
            }
            loc[2 * (size_t)q] = (uint32_t)w;
            loc[2 * (size_t)q + 1] = (uint32_t)(w >> 32);
        }
    }
    // halo nodes: their partial entries, in ascending patch order
    const int nentries = halo_base[npatch];
    std::vector<int> hcount(np, 0);
    for (int n : hnode_of_entry) ++hcount[n];
    std::vector<int> hnode, hptr(1, 0), hpos(np, -1);
    for (int n = 0; n < np; ++n)
        if (hcount[n] > 0) {
            hpos[n] = hptr.back();
            hnode.push_back(n + 1);
            hptr.push_back(hptr.back() + hcount[n]);
        }
    std::vector<int> hidx(nentries);
    for (int t = 0; t < nentries; ++t) hidx[hpos[hnode_of_entry[t]]++] = t;  // t ascends with p

    PatchHost h;
    h.elem_base.resize(npatch + 1);
    for (int p = 0; p <= npatch; ++p) h.elem_base[p] = std::min(ne, p * pe);
    h.npatch = npatch;
    h.max_nodes = max_nodes;
    h.nentries = nentries;
    h.perm = std::move(perm);
    h.node_base = std::move(node_base);
    h.n_int = std::move(n_int);
    h.halo_base = std::move(halo_base);
    h.pnode = std::move(pnode);
    h.lptr = std::move(lptr);
    h.loc = std::move(loc);
    h.hnode = std::move(hnode);
    h.hptr = std::move(hptr);
    h.hidx = std::move(hidx);
    return h;
}

void patch_numbering(const int* inpoel, const double* X, const double* Y, int ne, int np, int pe,
                     std::vector<int>& elem_old, std::vector<int>& node_old) {
    const PatchHost h = patches_host(inpoel, X, Y, ne, np, pe);
    elem_old = h.perm;
    node_old.clear();
    node_old.reserve(np);
    std::vector<char> seen(np, 0);
    for (int p = 0; p < h.npatch; ++p)  // each patch's nodes as listed: interior first
        for (int l = h.node_base[p]; l < h.node_base[p + 1]; ++l) {
            const int n = h.pnode[l] - 1;
            if (!seen[n]) {
                seen[n] = 1;
                node_old.push_back(n);
            }
        }
    for (int n = 0; n < np; ++n)  // nodes in no element keep their relative order at the end
        if (!seen[n]) node_old.push_back(n);
}

static void upload_patches(PatchSet& ps, const PatchHost& h, int ne, int pe, int block,
                           hipStream_t s) {
    const int npatch = h.npatch, nentries = h.nentries;
    std::vector<int> inv(ne);
    for (int q = 0; q < ne; ++q) inv[h.perm[q]] = q;
    ps.npatch = npatch;
    ps.pe = pe;
    ps.block = block;
    ps.max_nodes = h.max_nodes;
    ps.nentries = nentries;
    ps.nhalo = (int)h.hnode.size();
    ps.perm.alloc(ne);
    ps.perm.upload(h.perm.data(), ne, s);
    ps.inv_perm.alloc(ne);
    ps.inv_perm.upload(inv.data(), ne, s);
    ps.loc.alloc(2 * (size_t)ne);
    ps.loc.upload(h.loc.data(), 2 * (size_t)ne, s);
    ps.elem_base.alloc(npatch + 1);
    ps.elem_base.upload(h.elem_base.data(), npatch + 1, s);
    ps.node_base.alloc(npatch + 1);
    ps.node_base.upload(h.node_base.data(), npatch + 1, s);
    ps.n_int.alloc(npatch);
    ps.n_int.upload(h.n_int.data(), npatch, s);
    ps.halo_base.alloc(npatch + 1);
    ps.halo_base.upload(h.halo_base.data(), npatch + 1, s);
    ps.pnode.alloc(h.pnode.size());
    ps.pnode.upload(h.pnode.data(), h.pnode.size(), s);
    ps.lptr.alloc(h.lptr.size());
    ps.lptr.upload(h.lptr.data(), h.lptr.size(), s);
    ps.hnode.alloc(std::max<size_t>(1, h.hnode.size()));
    ps.hnode.upload(h.hnode.data(), h.hnode.size(), s);
    ps.hptr.alloc(h.hptr.size());
    ps.hptr.upload(h.hptr.data(), h.hptr.size(), s);
    ps.hidx.alloc(std::max(1, nentries));
    ps.hidx.upload(h.hidx.data(), nentries, s);
    ps.hpart.alloc(4 * (size_t)std::max(1, nentries));
    CFD_HIP(hipStreamSynchronize(s));  // the host vectors go out of scope
}

static void check_shape(int pe, int block) {
    // local slots are 11-bit and local node ids 10-bit fields of the per-element word
    CFD_CHECK(pe >= 1 && 3 * pe <= 2047 && kPatchMaxNodes <= 1023, "internal: patch size");
    CFD_CHECK(block >= 64 && block <= 512 && block % 64 == 0, "internal: patch block size");
}

void build_patches(PatchSet& ps, const int* inpoel, const double* X, const double* Y, int ne, int np,
                   int pe, int block, hipStream_t s) {
    check_shape(pe, block);
    const PatchHost h = patches_host(inpoel, X, Y, ne, np, pe);
    upload_patches(ps, h, ne, pe, block, s);
}

void build_patches_batch(PatchSet& ps, int ncases, const int* const* inpoel,
                         const double* const* X, const double* const* Y, const int* ne,
                         const int* np, int pe, int block, hipStream_t s) {
    check_shape(pe, block);
    // each case's patches as on its own, then concatenated: element, node,
    // patch, local-node and halo-entry numbers shifted by the cases before
    PatchHost u;
    std::vector<int> pcase, hcase;
    u.elem_base.push_back(0);
    u.node_base.push_back(0);
    u.halo_base.push_back(0);
    u.hptr.push_back(0);
    long long ne_tot = 0, np_tot = 0;
    for (int k = 0; k < ncases; ++k) {
        const PatchHost h = patches_host(inpoel[k], X[k], Y[k], ne[k], np[k], pe);
        const int E = (int)ne_tot, N = (int)np_tot;
        const int NB = u.node_base.back(), H = u.halo_base.back();
        for (int e : h.perm) u.perm.push_back(e + E);
        u.loc.insert(u.loc.end(), h.loc.begin(), h.loc.end());
        for (int p = 1; p <= h.npatch; ++p) {
            u.elem_base.push_back(h.elem_base[p] + E);
            u.node_base.push_back(h.node_base[p] + NB);
            u.halo_base.push_back(h.halo_base[p] + H);
        }
        u.n_int.insert(u.n_int.end(), h.n_int.begin(), h.n_int.end());
        for (int n : h.pnode) u.pnode.push_back(n + N);
        u.lptr.insert(u.lptr.end(), h.lptr.begin(), h.lptr.end());
        for (int n : h.hnode) u.hnode.push_back(n + N);
        for (size_t i = 1; i < h.hptr.size(); ++i) u.hptr.push_back(h.hptr[i] + H);
        for (int t : h.hidx) u.hidx.push_back(t + H);
        pcase.insert(pcase.end(), h.npatch, k);
        hcase.insert(hcase.end(), h.hnode.size(), k);
        u.npatch += h.npatch;
        u.nentries += h.nentries;
        u.max_nodes = std::max(u.max_nodes, h.max_nodes);
        ne_tot += ne[k];
        np_tot += np[k];
        CFD_CHECK(ne_tot < (1ll << 31) && np_tot < (1ll << 31), "arg: batch too large");
    }
    upload_patches(ps, u, (int)ne_tot, pe, block, s);
    ps.pcase.alloc(std::max<size_t>(1, pcase.size()));
    ps.pcase.upload(pcase.data(), pcase.size(), s);
    ps.hcase.alloc(std::max<size_t>(1, hcase.size()));
    ps.hcase.upload(hcase.data(), hcase.size(), s);
    CFD_HIP(hipStreamSynchronize(s));
}

}  // namespace cfd
