// msx_binning.hip -- binning of device-resident atoms for the potential splat.
//
// The splat kernel reads atoms grouped by (slice, lateral bin), each group in
// input order, with the start of every group (ModelPotential.cu:27-99 buckets
// the atoms of each slice on the host).  For atoms that already live in HBM
// (frozen-phonon configurations displaced on the GPU) the same grouping is
// built here without a round trip through the host: one pass computes each
// atom's cell key, a stable radix sort orders the (key, index) pairs, a
// lower-bound search per cell gives the cell starts and a gather writes the
// records.  Keys, records and cell starts equal those of the host binning.
#include <climits>

#include <hipcub/hipcub.hpp>

#include "msx_internal.h"

namespace msx {
namespace {

constexpr int kBinThreads = 256;

__global__ void __launch_bounds__(kBinThreads) bin_flags_kernel(int* flags) {
    const int t = threadIdx.x;
    if (t < 104) flags[t] = t == 0 ? INT_MAX : 0;
}

// key = cell of the atom, or ncell for an atom outside the slab or invalid
// (sorted after every cell); val = the atom's index
__global__ void __launch_bounds__(kBinThreads) bin_keys_kernel(DevBinArgs a) {
    const int i = blockIdx.x * kBinThreads + threadIdx.x;
    if (i >= a.n) return;
    const AtomIO at = a.atoms[i];
    a.val[i] = i;
    unsigned key = a.ncell;
    if (at.Z < 1 || at.Z > 103 || !isfinite(at.x) || !isfinite(at.y) || !isfinite(at.z)) {
        atomicMin(&a.flags[0], i);
    } else {
        a.flags[at.Z] = 1;
        long long cell;
        double xw, yw;
        if (bin_atom(a.g, at.x, at.y, at.z, &cell, &xw, &yw)) key = (unsigned)cell;
    }
    a.key[i] = key;
}

// cell_start[k] = first sorted position whose key is >= k, k = 0 .. ncell
// (cell_start[ncell] = number of atoms inside the slab)
__global__ void __launch_bounds__(kBinThreads) cell_start_kernel(const unsigned* __restrict__ keys, int n, int ncell,
                                                                 int* __restrict__ cell_start) {
    const long long stride = (long long)gridDim.x * kBinThreads;
    for (long long k = (long long)blockIdx.x * kBinThreads + threadIdx.x; k <= ncell; k += stride) {
        int lo = 0, hi = n;
        while (lo < hi) {
            const int mid = lo + ((hi - lo) >> 1);
            if ((long long)keys[mid] < k)
                lo = mid + 1;
            else
                hi = mid;
        }
        cell_start[k] = lo;
    }
}

// records of the atoms inside the slab, in sorted order
__global__ void __launch_bounds__(kBinThreads) gather_recs_kernel(const AtomIO* __restrict__ atoms,
                                                                  const int* __restrict__ order,
                                                                  const int* __restrict__ cell_start, int n,
                                                                  int ncell, BinGeom g, SlotMap slots,
                                                                  AtomRec* __restrict__ recs) {
    const int j = blockIdx.x * kBinThreads + threadIdx.x;
    if (j >= n || j >= cell_start[ncell]) return;
    const AtomIO at = atoms[order[j]];
    long long cell;
    double xw, yw;
    bin_atom(g, at.x, at.y, at.z, &cell, &xw, &yw);
    recs[j] = AtomRec{xw, yw, at.occ, (int16_t)at.Z, slots.s[at.Z]};
}

int end_bit(int ncell) {
    int b = 1;
    while (b < 31 && (1LL << b) <= ncell) b++;
    return b;
}

}  // namespace

hipError_t bin_sort_temp_bytes(int n, int ncell, size_t* bytes) {
    *bytes = 0;
    if (n <= 0) return hipSuccess;
    return hipcub::DeviceRadixSort::SortPairs(nullptr, *bytes, (const unsigned*)nullptr, (unsigned*)nullptr,
                                              (const int*)nullptr, (int*)nullptr, n, 0, end_bit(ncell));
}

hipError_t launch_bin_keys(const DevBinArgs& a, hipStream_t s) {
    bin_flags_kernel<<<1, kBinThreads, 0, s>>>(a.flags);
    if (a.n > 0) bin_keys_kernel<<<(a.n + kBinThreads - 1) / kBinThreads, kBinThreads, 0, s>>>(a);
    return hipGetLastError();
}

hipError_t launch_bin_sort(const DevBinArgs& a, int* cell_start, AtomRec* recs, const SlotMap& slots,
                           hipStream_t s) {
    if (a.n > 0) {
        size_t tb = a.temp_bytes;
        hipError_t e = hipcub::DeviceRadixSort::SortPairs(a.temp, tb, a.key, a.key2, a.val, a.val2, a.n, 0,
                                                          end_bit(a.ncell), s);
        if (e != hipSuccess) return e;
    }
    const long long cells = (long long)a.ncell + 1;
    const int grid = (int)std::min<long long>((cells + kBinThreads - 1) / kBinThreads, 8192);
    cell_start_kernel<<<grid, kBinThreads, 0, s>>>(a.key2, a.n, a.ncell, cell_start);
    if (a.n > 0)
        gather_recs_kernel<<<(a.n + kBinThreads - 1) / kBinThreads, kBinThreads, 0, s>>>(
            a.atoms, a.val2, cell_start, a.n, a.ncell, a.g, slots, recs);
    return hipGetLastError();
}

}  // namespace msx
