/* ----------------------------------------------------------------------
   LAMMPS pair style jax/mtp on the MI355X HIP evaluator (libmtp_hip.so).

   Flow per force call, as in the reference PairJaxMTP::compute
   (pair_jax_mtp.cpp:69-207):
     1. pack the full neighbor list into padded per-atom rows (the
        reference's loop at :96-173) and stage them (zero_overhead.cpp:61-117),
     2. evaluate on the GPU (replaces ZeroOverheadManager::execute_potential,
        zero_overhead.cpp:171-272),
     3. add forces for ii < inum, energy if eflag_global, virial if
        vflag_global.
   Steps 1-2 are one call, mtp_compute_full_list(): positions, types, tags
   and the full list go to the GPU as they are and a kernel packs and stages
   the rows there.  JAXMTP_HOST_PACK=1 (and triclinic boxes) keep the
   reference's split: mtp::pack_full_list on the host, then mtp_compute().
   Both give the same rows bit for bit.

   Differences, each deliberate:
     * the evaluation is FP64 (the reference cast r to float);
     * virial[] receives the LAMMPS-convention virial (-sum r (x) g, order
       xx yy zz xy xz yz, energy units); the reference added its
       volume-normalised stress in another order (pair_jax_mtp.cpp:203-205,
       README:40,47).  Set JAXMTP_VIRIAL_COMPAT=1 to get the old behaviour;
     * natoms > MAX_ATOMS and neighbor truncation are reported, not silent;
     * the cell volume of a triclinic box is xprd * yprd * zprd.
------------------------------------------------------------------------- */
#include "pair_jax_mtp.h"

#include <cstdlib>
#include <cstring>
#include <string>

#include "atom.h"
#include "comm.h"
#include "domain.h"
#include "error.h"
#include "force.h"
#include "memory.h"
#include "mtp_hip.h"
#include "nbr_pack.hpp"
#include "neigh_list.h"
#include "neighbor.h"
#include "update.h"

using namespace LAMMPS_NS;

PairJaxMTP::PairJaxMTP(LAMMPS *lmp) : Pair(lmp)
{
  single_enable = 0;
  restartinfo = 0;
  one_coeff = 1;
  manybody_flag = 1;
  mtp_path = nullptr;
  cutoff = 0.0;
  max_atoms = 0;
  max_neighbors = 0;
  debug_level = 0;
  ctx = nullptr;
  total_calls = 0;
  total_truncated = 0;
  const char *compat = std::getenv("JAXMTP_VIRIAL_COMPAT");
  virial_compat = (compat && std::strcmp(compat, "1") == 0) ? 1 : 0;
  const char *hp = std::getenv("JAXMTP_HOST_PACK");
  host_pack = (hp && std::strcmp(hp, "1") == 0) ? 1 : 0;
  if (comm->me == 0) utils::logmesg(lmp, "JAX-MTP (MI355X HIP evaluator) initialized\n");
}

PairJaxMTP::~PairJaxMTP()
{
  if (ctx) mtp_close(ctx);
  delete[] mtp_path;
  if (allocated) {
    memory->destroy(setflag);
    memory->destroy(cutsq);
  }
  if (total_calls > 0 && comm->me == 0) {
    utils::logmesg(lmp, "\n=== JAX-MTP SUMMARY ===\nTotal evaluations: {}\n", total_calls);
    if (total_truncated) utils::logmesg(lmp, "Neighbors dropped by MAX_ATOMS_PER_NEIGHBOR: {}\n", total_truncated);
  }
}

void PairJaxMTP::initialize_persistent_arrays()
{
  if (!pos.empty()) return;
  const size_t A = (size_t) max_atoms, K = (size_t) max_neighbors;
  pos.assign(A * 3 * K, 0.0);
  nbr.assign(A * K, 0);
  ntype.assign(A * K, 0);
  counts.assign(A, 0);
  atom_types.assign(A, -1);
  forces.assign(A * 3, 0.0);
  pos_rows.resize(A);
  nbr_rows.resize(A);
  ntype_rows.resize(A);
  for (size_t i = 0; i < A; ++i) {
    pos_rows[i] = pos.data() + i * 3 * K;
    nbr_rows[i] = nbr.data() + i * K;
    ntype_rows[i] = ntype.data() + i * K;
  }
}

void PairJaxMTP::compute(int eflag, int vflag)
{
  ev_init(eflag, vflag);
  initialize_persistent_arrays();

  double **x = atom->x;
  double **f = atom->f;
  NeighList *list = neighbor->lists[0];
  const int inum = list->inum;
  const int nlocal = atom->nlocal;
  if (nlocal > max_atoms)
    error->all(FLERR, "jax/mtp: {} local atoms exceed MAX_ATOMS = {}", nlocal, max_atoms);

  // LAMMPS's h = (xprd, yprd, zprd, yz, xz, xy) is upper triangular, so the
  // cell volume is xprd * yprd * zprd for triclinic boxes too.  (The
  // reference, pair_jax_mtp.cpp:179-181, took a 3x3 determinant of h[0..8],
  // reading past the six elements of Domain::h for triclinic boxes.)
  const double volume = domain->xprd * domain->yprd * domain->zprd;

  double energy = 0.0, stress6[6], virial6[6];
  long long truncated = 0;
  int rc;
  if (host_pack || domain->triclinic) {
    // the reference's flow: rows packed on the host, then staged and evaluated
    auto min_image = [this](double &dx, double &dy, double &dz) {
      domain->minimum_image(FLERR, dx, dy, dz);
    };
    mtp::PackOut out{pos.data(), nbr.data(), ntype.data(), counts.data(), atom_types.data()};
    const long long nneigh_total =
        mtp::pack_full_list(nlocal, x, atom->type, atom->tag, inum, list->ilist, list->numneigh, list->firstneigh,
                            cutoff, max_atoms, max_neighbors, min_image, out, tag_to_dense, &truncated);
    rc = mtp_compute(ctx, nlocal, (int) nneigh_total, pos_rows.data(), atom_types.data(), nbr_rows.data(),
                     counts.data(), ntype_rows.data(), volume, &energy, forces.data(), stress6, virial6);
  } else {
    // the same rows, packed and staged on the GPU from the uploaded full list
    const double prd[3] = {domain->xprd, domain->yprd, domain->zprd};
    // the list, types and tags are re-sent only after a neighbor-list rebuild (neighbor->ago == 0)
    rc = mtp_compute_full_list(ctx, nlocal, nlocal + atom->nghost, x[0], atom->type, atom->tag, inum, list->ilist,
                               list->numneigh, list->firstneigh, neighbor->ago == 0 ? 1 : 0, prd,
                               domain->periodicity, volume, &energy, forces.data(), stress6, virial6, &truncated);
  }
  if (rc != MTP_OK) error->all(FLERR, "jax/mtp evaluation failed: {}", mtp_last_error(ctx));
  total_truncated += truncated;
  if (truncated && debug_level > 0)
    error->warning(FLERR, "jax/mtp: {} in-cutoff neighbors dropped (MAX_ATOMS_PER_NEIGHBOR = {})", truncated,
                   max_neighbors);

  for (int ii = 0; ii < inum; ii++) {
    const int i = list->ilist[ii];
    f[i][0] += forces[3 * i + 0];
    f[i][1] += forces[3 * i + 1];
    f[i][2] += forces[3 * i + 2];
  }
  if (eflag_global) eng_vdwl += energy;
  if (vflag_global) {
    const double *v = virial_compat ? stress6 : virial6;
    for (int k = 0; k < 6; k++) virial[k] += v[k];
  }
  // per-atom energy and virial (compute pe/atom, stress/atom): E_i and the symmetrised -W_i of
  // atom i's row, which sum to the totals above (the reference has no per-atom values)
  if (eflag_atom || vflag_atom) {
    atom_e.resize(nlocal);
    atom_v.resize((size_t) nlocal * 6);
    if (mtp_last_atom_values(ctx, nlocal, eflag_atom ? atom_e.data() : nullptr,
                             vflag_atom ? atom_v.data() : nullptr) != MTP_OK)
      error->all(FLERR, "jax/mtp per-atom values failed: {}", mtp_last_error(ctx));
    for (int ii = 0; ii < inum; ii++) {
      const int i = list->ilist[ii];
      if (eflag_atom) eatom[i] += atom_e[i];
      if (vflag_atom)
        for (int k = 0; k < 6; k++) vatom[i][k] += atom_v[6 * (size_t) i + k];
    }
  }
  total_calls++;
}

void PairJaxMTP::settings(int narg, char **arg)
{
  if (narg < 4)
    error->all(FLERR, "Illegal pair_style command - usage: pair_style jax/mtp <file.mtp> <max_atoms> "
                      "<max_neighbors> <cutoff> [debug_level]");
  delete[] mtp_path;
  mtp_path = utils::strdup(arg[0]);
  max_atoms = utils::inumeric(FLERR, arg[1], false, lmp);
  max_neighbors = utils::inumeric(FLERR, arg[2], false, lmp);
  cutoff = utils::numeric(FLERR, arg[3], false, lmp);
  if (narg >= 5) debug_level = utils::inumeric(FLERR, arg[4], false, lmp);
  if (max_atoms <= 0 || max_neighbors <= 0 || cutoff <= 0.0)
    error->all(FLERR, "max_atoms, max_neighbors, and cutoff must be positive");
}

void PairJaxMTP::coeff(int narg, char **arg)
{
  if (narg != 2) error->all(FLERR, "Incorrect args for pair coefficients");
  if (!allocated) allocate();
  for (int i = 1; i <= atom->ntypes; i++)
    for (int j = 1; j <= atom->ntypes; j++) {
      setflag[i][j] = 1;
      cutsq[i][j] = cutoff * cutoff;
    }
}

double PairJaxMTP::init_one(int i, int j)
{
  if (setflag[i][j] == 0) error->all(FLERR, "All pair coeffs are not set");
  return cutoff;
}

void PairJaxMTP::init_style()
{
  if (atom->tag_enable == 0) error->all(FLERR, "Pair style jax/mtp requires atom IDs");
  if (max_atoms <= 0 || max_neighbors <= 1)
    error->all(FLERR, "System capacity not set - check pair_style arguments");
  if (comm->nprocs > 1)
    error->all(FLERR, "Pair style jax/mtp folds ghosts onto owners by tag and runs on one MPI rank");
  neighbor->add_request(this, NeighConst::REQ_FULL);
  if (!ctx) {
    const int rc = mtp_open(mtp_path, max_atoms, max_neighbors, cutoff, debug_level, &ctx);
    if (rc != MTP_OK) error->all(FLERR, "jax/mtp: cannot open potential: {}", mtp_global_error());
  }
}

void PairJaxMTP::allocate()
{
  allocated = 1;
  const int n = atom->ntypes;
  memory->create(setflag, n + 1, n + 1, "pair:setflag");
  memory->create(cutsq, n + 1, n + 1, "pair:cutsq");
}
