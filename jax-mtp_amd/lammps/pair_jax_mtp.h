/* ----------------------------------------------------------------------
   LAMMPS pair style jax/mtp, re-targeted to the MI355X HIP evaluator.

   Same class name, style name, arguments and checks as the reference
   (pair_style_files/pair_jax_mtp.h:6-51, pair_jax_mtp.cpp:228-282):

     pair_style jax/mtp <file.mtp> MAX_ATOMS MAX_ATOMS_PER_NEIGHBOR MTP_CUTOFF [DEBUG_LEVEL]
     pair_coeff * *

   The first argument is now a trained .mtp potential (the reference took a
   jax.export .bin).  Evaluation goes through the C ABI in mtp_hip.h; no
   Python is involved.
------------------------------------------------------------------------- */
#ifdef PAIR_CLASS
// clang-format off
PairStyle(jax/mtp, PairJaxMTP)
// clang-format on
#else
#ifndef LMP_PAIR_JAX_MTP_H
#define LMP_PAIR_JAX_MTP_H

#include <vector>

#include "pair.h"

struct mtp_ctx;

namespace LAMMPS_NS {

class PairJaxMTP : public Pair {
 public:
  PairJaxMTP(class LAMMPS *);
  ~PairJaxMTP() override;

  void compute(int, int) override;
  void settings(int, char **) override;
  void coeff(int, char **) override;
  double init_one(int, int) override;
  void init_style() override;

 private:
  char *mtp_path;
  double cutoff;
  int max_atoms, max_neighbors;
  int debug_level;
  int virial_compat;  // 1: add the reference's stress6 into virial[] (JAXMTP_VIRIAL_COMPAT=1)
  int host_pack;      // 1: pack the rows on the host (JAXMTP_HOST_PACK=1); 0: on the GPU
  mtp_ctx *ctx;
  long long total_calls, total_truncated;

  std::vector<double> pos;        // [max_atoms][3*max_neighbors]
  std::vector<int> nbr, ntype;    // [max_atoms][max_neighbors]
  std::vector<int> counts, atom_types, tag_to_dense;
  std::vector<const double *> pos_rows;
  std::vector<const int *> nbr_rows, ntype_rows;
  std::vector<double> forces;     // [max_atoms][3]
  std::vector<double> atom_e, atom_v;  // per-atom energy [nlocal], virial [nlocal][6]

  void allocate();
  void initialize_persistent_arrays();
};

}  // namespace LAMMPS_NS
#endif
#endif
