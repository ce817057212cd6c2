// GPU neighbor packing + staging (mtp_pack.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace mtp {

// Device arrays of one packing pass.  The full neighbor list is flattened
// in ilist order: the neighbors of ilist[ii] are nbr[off[ii] .. off[ii+1]).
struct PackArgs {
    int nlocal, inum, max_nbr, max_tag;
    double cutoff;
    double prd[3];
    int periodic[3];
    const double* x;       // [nall][3]
    const int* type;       // [nall], 1-based LAMMPS types
    const int* tag;        // [nall]
    const int* ilist;      // [inum]
    const long long* off;  // [inum + 1]
    const int* nbr;        // [off[inum]]
    int* map;              // [max_tag + 1] scratch: tag -> local index
    int* itypes;           // [nlocal]            staged rows (the device contract of the evaluation)
    int* js;               // [nlocal][max_nbr]
    int* jtypes;           // [nlocal][max_nbr]
    double* rij;           // [nlocal][max_nbr][3]
    unsigned long long* counters;  // [2]: neighbors kept, in-cutoff neighbors dropped (row full)
};

hipError_t launch_pack(const PackArgs& a, hipStream_t stream);

}  // namespace mtp
