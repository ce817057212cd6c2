// Neighbor packing and staging on the GPU: the pair style's full-list walk
// (PairJaxMTP::compute, pair_jax_mtp.cpp:96-173) and the staging of
// ZeroOverheadManager (zero_overhead.cpp:61-117) in one pass, from the
// positions, types, tags and full neighbor list of LAMMPS uploaded as they
// are (the list flattened in ilist order).
//
// Results are bit for bit those of the host path (mtp::pack_full_list then
// mtp_stage_host): the same minimum image, the same r^2 = (dx dx + dy dy) +
// dz dz without fused multiply-adds, the same first max_nbr in-cutoff
// neighbors in list order, the same tag fold of ghosts onto owners
// (unmapped -> nlocal, type -1), and the staging sentinels of unused slots
// (j = nlocal, type = -1, r = 2000).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mtp_pack.hpp"
#include "mtp_plan.hpp"

namespace mtp {

namespace {

constexpr int kNeighMask = 0x1FFFFFFF;  // LAMMPS NEIGHMASK

// Rows of every local atom: type - 1 and all slots empty; tag -> local
// index for the tags of the locals (map pre-filled with -1); counters cleared.
__global__ void pack_fill_kernel(int nlocal, int max_nbr, const int* type, const int* tag, int max_tag, int* map,
                                 int* itypes, int* js, int* jtypes, double* rij, unsigned long long* counters)
{
    const long long total = (long long)nlocal * max_nbr;
    const long long stride = (long long)gridDim.x * blockDim.x;
    const long long t0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    for (long long s = t0; s < total; s += stride) {
        js[s] = nlocal;
        jtypes[s] = -1;
        rij[3 * s] = 2000.0;
        rij[3 * s + 1] = 2000.0;
        rij[3 * s + 2] = 2000.0;
    }
    for (long long i = t0; i < nlocal; i += stride) {
        itypes[i] = type[i] - 1;
        const int tg = tag[i];
        if (tg >= 0 && tg <= max_tag) map[tg] = (int)i;
    }
    if (t0 == 0) { counters[0] = 0ull; counters[1] = 0ull; }
}

__global__ void pack_map_clear_kernel(int* map, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) map[i] = -1;
}

// One wavefront per list entry ii: the row of atom i = ilist[ii] from its
// neighbors nbr[off[ii] .. off[ii + 1]), 64 at a time; a ballot keeps the
// list order of the in-cutoff ones.
#pragma clang fp contract(off)
__global__ void __launch_bounds__(256) pack_rows_kernel(PackArgs a)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int ii = (int)(((long long)blockIdx.x * blockDim.x + threadIdx.x) / kWave);
    if (ii >= a.inum) return;
    const int i = a.ilist[ii];
    const double xi = a.x[3 * (size_t)i], yi = a.x[3 * (size_t)i + 1], zi = a.x[3 * (size_t)i + 2];
    const double cutsq = a.cutoff * a.cutoff;
    const long long b0 = a.off[ii], b1 = a.off[ii + 1];
    int* jrow = a.js + (size_t)i * a.max_nbr;
    int* trow = a.jtypes + (size_t)i * a.max_nbr;
    double* rrow = a.rij + (size_t)i * a.max_nbr * 3;
    int cnt = 0;            // in-cutoff neighbors seen (wave-uniform)
    for (long long c = b0; c < b1; c += kWave) {
        const long long jj = c + lane;
        bool in = false;
        double dx = 0.0, dy = 0.0, dz = 0.0;
        int j = 0;
        if (jj < b1) {
            j = a.nbr[jj] & kNeighMask;
            if (j != i) {
                dx = a.x[3 * (size_t)j] - xi;
                dy = a.x[3 * (size_t)j + 1] - yi;
                dz = a.x[3 * (size_t)j + 2] - zi;
                // Domain::minimum_image, orthogonal box
                if (a.periodic[0] && (dx < 0 ? -dx : dx) > 0.5 * a.prd[0]) dx += dx < 0.0 ? a.prd[0] : -a.prd[0];
                if (a.periodic[1] && (dy < 0 ? -dy : dy) > 0.5 * a.prd[1]) dy += dy < 0.0 ? a.prd[1] : -a.prd[1];
                if (a.periodic[2] && (dz < 0 ? -dz : dz) > 0.5 * a.prd[2]) dz += dz < 0.0 ? a.prd[2] : -a.prd[2];
                const double rsq = dx * dx + dy * dy + dz * dz;
                in = rsq < cutsq;
            }
        }
        const uint64_t mask = __ballot(in);
        const int pos = cnt + __popcll(mask & ((1ull << lane) - 1ull));
        if (in && pos < a.max_nbr) {
            int dense = a.nlocal;
            const int tg = a.tag[j];
            if (tg >= 0 && tg <= a.max_tag && a.map[tg] != -1) dense = a.map[tg];
            jrow[pos] = dense;
            trow[pos] = dense != a.nlocal ? a.type[j] - 1 : -1;
            rrow[3 * pos] = dx;
            rrow[3 * pos + 1] = dy;
            rrow[3 * pos + 2] = dz;
        }
        cnt += __popcll(mask);
    }
    if (lane == 0) {
        const int kept = cnt < a.max_nbr ? cnt : a.max_nbr;
        atomicAdd(a.counters, (unsigned long long)kept);
        if (cnt > kept) atomicAdd(a.counters + 1, (unsigned long long)(cnt - kept));
    }
}

}  // namespace

hipError_t launch_pack(const PackArgs& a, hipStream_t stream)
{
    const int map_n = a.max_tag + 1;
    hipLaunchKernelGGL(pack_map_clear_kernel, dim3(map_n / 256 + 1 < 2048 ? map_n / 256 + 1 : 2048), dim3(256), 0,
                       stream, a.map, map_n);
    const long long total = (long long)a.nlocal * a.max_nbr;
    int grid = (int)((total + 255) / 256);
    if (grid > 4096) grid = 4096;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(pack_fill_kernel, dim3(grid), dim3(256), 0, stream, a.nlocal, a.max_nbr, a.type, a.tag,
                       a.max_tag, a.map, a.itypes, a.js, a.jtypes, a.rij, a.counters);
    if (a.inum > 0)
        hipLaunchKernelGGL(pack_rows_kernel, dim3((a.inum + 3) / 4), dim3(256), 0, stream, a);
    return hipGetLastError();
}

}  // namespace mtp
