// Mesh tables on the device: elements-surrounding-points (esup), points-
// surrounding-points (psup), the mesh-Laplace CSR structure, and the slot maps
// used by the deterministic gathers.
//
// The reference builds esup/psup serially with Loehner's linked lists
// (pointNeighbor.f90:5-91) and the CSR structure from psup
// (mLaplace.f90:60-94).  Here the same tables come from data-parallel passes
// whose results are identical entry for entry:
//   esup1: each node's element list in ascending element order (what the
//          serial fill over ascending elements produces) -- histogram with
//          atomics, scan, unordered fill, then a per-node sort of the segment;
//   psup1: each node's neighbours in first-discovery order over its esup
//          elements and their local nodes 1..3 -- one thread per node walks
//          its own esup segment exactly as the serial loop does;
//   lap_rowptr/lap_idx: diagonal first, then the psup row (mLaplace.f90:82-93).
#include <hipcub/hipcub.hpp>

#include "cfd_internal.hpp"

namespace cfd {
namespace {

constexpr int kMaxDegree = 64;  // neighbours per node accepted by the psup pass

__global__ void k_check_inpoel(const int* __restrict__ inpoel, int nelem, int npoin,
                               int* __restrict__ bad) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nelem) return;
    int a = inpoel[3 * e], b = inpoel[3 * e + 1], c = inpoel[3 * e + 2];
    if (a < 1 || a > npoin || b < 1 || b > npoin || c < 1 || c > npoin) atomicOr(bad, 1);
    else if (a == b || b == c || a == c) atomicOr(bad, 2);
}

__global__ void k_esup_count(const int* __restrict__ inpoel, long long n3, int* __restrict__ cnt) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n3) return;
    atomicAdd(&cnt[inpoel[t] - 1], 1);
}

__global__ void k_esup_fill(const int* __restrict__ inpoel, long long n3, int* __restrict__ cursor,
                            int* __restrict__ esup1) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n3) return;
    int node = inpoel[t] - 1;
    int pos = atomicAdd(&cursor[node], 1);
    esup1[pos] = (int)(t / 3) + 1;
}

// Sort each node's segment ascending (segments hold a handful of entries) and
// record, for every element, where it sits in each of its nodes' segments.
__global__ void k_esup_sort_epos(const int* __restrict__ inpoel, const int* __restrict__ esup2,
                                 int* __restrict__ esup1, int* __restrict__ epos, int npoin) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npoin) return;
    int b = esup2[i], e = esup2[i + 1];
    for (int s = b + 1; s < e; ++s) {
        int v = esup1[s];
        int j = s - 1;
        while (j >= b && esup1[j] > v) {
            esup1[j + 1] = esup1[j];
            --j;
        }
        esup1[j + 1] = v;
    }
    for (int s = b; s < e; ++s) {
        int el = esup1[s] - 1;
        for (int k = 0; k < 3; ++k)
            if (inpoel[3 * el + k] == i + 1) epos[3 * el + k] = s;
    }
}

// psup, count (fill == nullptr) or fill pass; discovery order as in
// pointNeighbor.f90:57-88.
__global__ void k_psup(const int* __restrict__ inpoel, const int* __restrict__ esup1,
                       const int* __restrict__ esup2, int* __restrict__ cnt,
                       const int* __restrict__ psup2, int* __restrict__ psup1, int npoin,
                       int* __restrict__ bad) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npoin) return;
    int seen[kMaxDegree];
    int n = 0;
    const int ip = i + 1;
    for (int s = esup2[i]; s < esup2[i + 1]; ++s) {
        int el = esup1[s] - 1;
        for (int k = 0; k < 3; ++k) {
            int jp = inpoel[3 * el + k];
            if (jp == ip) continue;
            bool dup = false;
            for (int q = 0; q < n; ++q) dup |= (seen[q] == jp);
            if (dup) continue;
            if (n == kMaxDegree) {
                atomicOr(bad, 4);
                return;
            }
            seen[n++] = jp;
        }
    }
    if (psup1 == nullptr) {
        cnt[i] = n;
    } else {
        int base = psup2[i];
        for (int q = 0; q < n; ++q) psup1[base + q] = seen[q];
    }
}

__global__ void k_lap_structure(const int* __restrict__ psup1, const int* __restrict__ psup2,
                                int* __restrict__ rowptr, int* __restrict__ idx, int npoin) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > npoin) return;
    rowptr[i] = psup2[i] + i;  // lap_rowptr(i) = psup2(i) + i - 1 (1-based)
    if (i == npoin) return;
    int r = psup2[i] + i;
    idx[r] = i + 1;
    int n = psup2[i + 1] - psup2[i];
    for (int q = 0; q < n; ++q) idx[r + 1 + q] = psup1[psup2[i] + q];
}

// For esup entry s (node i, element e): CSR positions in row i of inpoel(1..3,e).
__global__ void k_lap_slots(const int* __restrict__ inpoel, const int* __restrict__ esup1,
                            const int* __restrict__ esup2, const int* __restrict__ rowptr,
                            const int* __restrict__ idx, int* __restrict__ lslot, int npoin,
                            int* __restrict__ bad) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npoin) return;
    int rb = rowptr[i], re = rowptr[i + 1];
    for (int s = esup2[i]; s < esup2[i + 1]; ++s) {
        int el = esup1[s] - 1;
        for (int j = 0; j < 3; ++j) {
            int col = inpoel[3 * el + j];
            int pos = -1;
            for (int k = rb; k < re; ++k)
                if (idx[k] == col) { pos = k; break; }
            if (pos < 0) atomicOr(bad, 8);
            lslot[3 * s + j] = pos;
        }
    }
}

void inclusive_scan_to(const int* in, int* out_plus1, int n, hipStream_t s) {
    size_t bytes = 0;
    CFD_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, bytes, in, out_plus1, n, s));
    DBuf<unsigned char> tmp(bytes ? bytes : 1);
    CFD_HIP(hipcub::DeviceScan::InclusiveSum(tmp.p, bytes, in, out_plus1, n, s));
    CFD_HIP(hipStreamSynchronize(s));
}

int read_flag(DBuf<int>& flag, hipStream_t s) {
    int h = 0;
    flag.download(&h, 1, s);
    CFD_HIP(hipStreamSynchronize(s));
    return h;
}

}  // namespace

void build_tables(Tables& t, const int* inpoel_host, int nelem, int npoin, hipStream_t s) {
    CFD_CHECK(nelem > 0 && npoin > 0, "empty mesh");
    CFD_CHECK((long long)nelem * 9 < (1LL << 31), "mesh too large for int32 slot maps");
    t.nelem = nelem;
    t.npoin = npoin;
    const long long n3 = 3LL * nelem;
    t.inpoel.alloc(n3);
    t.inpoel.upload(inpoel_host, n3, s);

    DBuf<int> bad(1);
    bad.zero(s);
    k_check_inpoel<<<grid_for(nelem), kBlock, 0, s>>>(t.inpoel.p, nelem, npoin, bad.p);
    int b = read_flag(bad, s);
    if (b & 1) throw Error("mesh: node id out of range in inpoel");
    if (b & 2) throw Error("mesh: element with a repeated node");

    // esup
    DBuf<int> cnt(npoin);
    cnt.zero(s);
    k_esup_count<<<grid_for(n3), kBlock, 0, s>>>(t.inpoel.p, n3, cnt.p);
    t.esup2.alloc(npoin + 1);
    t.esup2.zero(s);
    inclusive_scan_to(cnt.p, t.esup2.p + 1, npoin, s);
    CFD_HIP(hipMemcpyAsync(cnt.p, t.esup2.p, sizeof(int) * npoin, hipMemcpyDeviceToDevice, s));
    t.esup1.alloc(n3);
    k_esup_fill<<<grid_for(n3), kBlock, 0, s>>>(t.inpoel.p, n3, cnt.p, t.esup1.p);
    t.epos.alloc(n3);
    k_esup_sort_epos<<<grid_for(npoin), kBlock, 0, s>>>(t.inpoel.p, t.esup2.p, t.esup1.p, t.epos.p,
                                                        npoin);

    // psup
    k_psup<<<grid_for(npoin), kBlock, 0, s>>>(t.inpoel.p, t.esup1.p, t.esup2.p, cnt.p, nullptr,
                                              nullptr, npoin, bad.p);
    t.psup2.alloc(npoin + 1);
    t.psup2.zero(s);
    inclusive_scan_to(cnt.p, t.psup2.p + 1, npoin, s);
    if (read_flag(bad, s) & 4) throw Error("mesh: node with more than 64 neighbours");
    int npsup = 0;
    CFD_HIP(hipMemcpyAsync(&npsup, t.psup2.p + npoin, sizeof(int), hipMemcpyDeviceToHost, s));
    CFD_HIP(hipStreamSynchronize(s));
    t.npsup = npsup;
    t.psup1.alloc(npsup ? npsup : 1);
    k_psup<<<grid_for(npoin), kBlock, 0, s>>>(t.inpoel.p, t.esup1.p, t.esup2.p, cnt.p, t.psup2.p,
                                              t.psup1.p, npoin, bad.p);

    // Laplace CSR structure and slot map
    CFD_CHECK((long long)npsup + npoin < (1LL << 31), "CSR too large for int32");
    t.nnz = npsup + npoin;
    t.rowptr.alloc(npoin + 1);
    t.idx.alloc(t.nnz);
    k_lap_structure<<<grid_for(npoin + 1), kBlock, 0, s>>>(t.psup1.p, t.psup2.p, t.rowptr.p,
                                                            t.idx.p, npoin);
    t.lslot.alloc(3 * n3);
    k_lap_slots<<<grid_for(npoin), kBlock, 0, s>>>(t.inpoel.p, t.esup1.p, t.esup2.p, t.rowptr.p,
                                                   t.idx.p, t.lslot.p, npoin, bad.p);
    CFD_HIP(hipGetLastError());
    if (read_flag(bad, s) & 8) throw Error("mesh: Laplace slot map incomplete");
}

void reorder_esup(Tables& t, const std::vector<int>& key, hipStream_t s) {
    const int np = t.npoin, ne = t.nelem;
    CFD_CHECK((int)key.size() == ne, "internal: esup key size");
    std::vector<int> e1(3 * (size_t)ne), e2(np + 1), inp(3 * (size_t)ne), epos(3 * (size_t)ne);
    t.esup1.download(e1.data(), e1.size(), s);
    t.esup2.download(e2.data(), e2.size(), s);
    t.inpoel.download(inp.data(), inp.size(), s);
    CFD_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < np; ++i) {
        std::sort(e1.begin() + e2[i], e1.begin() + e2[i + 1],
                  [&](int a, int b) { return key[a - 1] < key[b - 1]; });
        for (int q = e2[i]; q < e2[i + 1]; ++q) {
            const int e = e1[q] - 1;
            for (int k = 0; k < 3; ++k)
                if (inp[3 * (size_t)e + k] == i + 1) epos[3 * (size_t)e + k] = q;
        }
    }
    t.esup1.upload(e1.data(), e1.size(), s);
    t.epos.upload(epos.data(), epos.size(), s);
    DBuf<int> bad(1);
    bad.zero(s);
    k_lap_slots<<<grid_for(np), kBlock, 0, s>>>(t.inpoel.p, t.esup1.p, t.esup2.p, t.rowptr.p,
                                                t.idx.p, t.lslot.p, np, bad.p);
    CFD_HIP(hipGetLastError());
    if (read_flag(bad, s) & 8) throw Error("mesh: Laplace slot map incomplete");
}

}  // namespace cfd
