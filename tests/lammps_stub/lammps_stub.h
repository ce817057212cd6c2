// Minimal stand-ins for the LAMMPS classes that pair_style jax/mtp
// (jax-mtp_amd/lammps/pair_jax_mtp.{h,cpp}) uses, so that the pair style's
// own source compiles and runs without LAMMPS (test infrastructure only).
// Names, members and call signatures follow the LAMMPS API the pair style
// calls: Pointers, Pair (ev_init, eng_vdwl, virial, eatom, vatom, setflag,
// cutsq), Atom, Comm, Domain (minimum_image with FLERR), Error (all,
// warning with "{}" formatting), Memory (create/destroy of 2-d arrays),
// Neighbor (lists, add_request), NeighList, Update and utils.  Errors throw
// std::runtime_error instead of aborting MPI.
#pragma once
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#define FLERR __FILE__, __LINE__

namespace LAMMPS_NS {

typedef int tagint;

namespace fmtlite {
inline void put(std::ostringstream& os, const char* f)
{
    os << f;
}
template <typename T, typename... R>
void put(std::ostringstream& os, const char* f, const T& v, const R&... rest)
{
    const char* p = std::strstr(f, "{}");
    if (!p) { os << f; return; }
    os.write(f, p - f);
    os << v;
    put(os, p + 2, rest...);
}
template <typename... A>
std::string format(const std::string& f, const A&... a)
{
    std::ostringstream os;
    put(os, f.c_str(), a...);
    return os.str();
}
}  // namespace fmtlite

class LAMMPS;

class Error {
 public:
    std::vector<std::string> warnings;
    template <typename... A>
    [[noreturn]] void all(const char* file, int line, const std::string& f, const A&... a)
    {
        throw std::runtime_error(fmtlite::format(f, a...) + " (" + file + ":" + std::to_string(line) + ")");
    }
    template <typename... A>
    void warning(const char*, int, const std::string& f, const A&... a)
    {
        warnings.push_back(fmtlite::format(f, a...));
    }
};

class Memory {
 public:
    template <typename T>
    T** create(T**& a, int n1, int n2, const char*)
    {
        T* data = new T[(size_t)n1 * n2]();
        a = new T*[n1];
        for (int i = 0; i < n1; ++i) a[i] = data + (size_t)i * n2;
        return a;
    }
    template <typename T>
    void destroy(T**& a)
    {
        if (a) { delete[] a[0]; delete[] a; }
        a = nullptr;
    }
};

class Atom {
 public:
    int nlocal = 0, nghost = 0, ntypes = 0, tag_enable = 1;
    double** x = nullptr;
    double** f = nullptr;
    int* type = nullptr;
    tagint* tag = nullptr;
};

class Comm {
 public:
    int me = 0, nprocs = 1;
};

class Domain {
 public:
    int triclinic = 0;
    int periodicity[3] = {1, 1, 1};
    double xprd = 0, yprd = 0, zprd = 0;
    double h[6] = {0, 0, 0, 0, 0, 0};  // xprd, yprd, zprd, yz, xz, xy
    // orthogonal periodic box (Domain::minimum_image for triclinic == 0)
    void minimum_image(const char*, int, double& dx, double& dy, double& dz) const
    {
        const double prd[3] = {xprd, yprd, zprd};
        double* d[3] = {&dx, &dy, &dz};
        for (int q = 0; q < 3; ++q) {
            if (!periodicity[q]) continue;
            if (std::fabs(*d[q]) > 0.5 * prd[q]) {
                if (*d[q] < 0.0) *d[q] += prd[q];
                else *d[q] -= prd[q];
            }
        }
    }
};

class NeighList {
 public:
    int inum = 0;
    int* ilist = nullptr;
    int* numneigh = nullptr;
    int** firstneigh = nullptr;
};

namespace NeighConst {
enum { REQ_FULL = 1 << 0 };
}

class Neighbor {
 public:
    NeighList* lists[1] = {nullptr};
    int ago = 0;  // steps since the last neighbor-list build (0: built for this step)
    int requests = 0, request_flags = 0;
    void add_request(void*, int flags)
    {
        ++requests;
        request_flags = flags;
    }
};

class Update {
 public:
    long ntimestep = 0;
};

class Force {};

class LAMMPS {
 public:
    Atom* atom = nullptr;
    Comm* comm = nullptr;
    Domain* domain = nullptr;
    Error* error = nullptr;
    Force* force = nullptr;
    Memory* memory = nullptr;
    Neighbor* neighbor = nullptr;
    Update* update = nullptr;
    std::string log;
};

class Pointers {
 public:
    explicit Pointers(LAMMPS* ptr)
        : lmp(ptr), atom(ptr->atom), comm(ptr->comm), domain(ptr->domain), error(ptr->error), force(ptr->force),
          memory(ptr->memory), neighbor(ptr->neighbor), update(ptr->update)
    {
    }
    virtual ~Pointers() = default;

 protected:
    LAMMPS* lmp;
    Atom*& atom;
    Comm*& comm;
    Domain*& domain;
    Error*& error;
    Force*& force;
    Memory*& memory;
    Neighbor*& neighbor;
    Update*& update;
};

namespace utils {
template <typename... A>
void logmesg(LAMMPS* lmp, const std::string& f, const A&... a)
{
    lmp->log += fmtlite::format(f, a...);
}
inline char* strdup(const std::string& s)
{
    char* p = new char[s.size() + 1];
    std::memcpy(p, s.c_str(), s.size() + 1);
    return p;
}
inline int inumeric(const char* file, int line, const char* s, bool, LAMMPS* lmp)
{
    char* end = nullptr;
    const long v = std::strtol(s, &end, 10);
    if (!s[0] || *end) lmp->error->all(file, line, "Expected integer parameter instead of '{}'", s);
    return (int)v;
}
inline double numeric(const char* file, int line, const char* s, bool, LAMMPS* lmp)
{
    char* end = nullptr;
    const double v = std::strtod(s, &end);
    if (!s[0] || *end) lmp->error->all(file, line, "Expected floating point parameter instead of '{}'", s);
    return v;
}
}  // namespace utils

// Pair: the members pair styles use for flags, accumulation and coefficients.
class Pair : protected Pointers {
 public:
    explicit Pair(LAMMPS* l) : Pointers(l) {}
    ~Pair() override
    {
        delete[] eatom;
        memory->destroy(vatom);
    }
    virtual void compute(int, int) = 0;
    virtual void settings(int, char**) = 0;
    virtual void coeff(int, char**) = 0;
    virtual double init_one(int, int) { return 0.0; }
    virtual void init_style() {}

    double eng_vdwl = 0.0, eng_coul = 0.0;
    double virial[6] = {0, 0, 0, 0, 0, 0};
    double* eatom = nullptr;
    double** vatom = nullptr;
    int eflag_either = 0, eflag_global = 0, eflag_atom = 0;
    int vflag_either = 0, vflag_global = 0, vflag_atom = 0;
    int single_enable = 1, restartinfo = 1, one_coeff = 0, manybody_flag = 0;
    int allocated = 0;
    int** setflag = nullptr;
    double** cutsq = nullptr;

 protected:
    int maxeatom = 0, maxvatom = 0;
    // eflag: 1 global energy, 2 per-atom energy; vflag: 1/2 global virial, 4 per-atom virial
    void ev_init(int eflag, int vflag)
    {
        eflag_global = eflag & 1;
        eflag_atom = (eflag & 2) ? 1 : 0;
        vflag_global = (vflag & 3) ? 1 : 0;
        vflag_atom = (vflag & 4) ? 1 : 0;
        eflag_either = eflag_global || eflag_atom;
        vflag_either = vflag_global || vflag_atom;
        eng_vdwl = eng_coul = 0.0;
        for (double& v : virial) v = 0.0;
        const int nall = atom->nlocal + atom->nghost;
        if (eflag_atom) {
            if (nall > maxeatom) { delete[] eatom; eatom = new double[nall]; maxeatom = nall; }
            for (int i = 0; i < nall; ++i) eatom[i] = 0.0;
        }
        if (vflag_atom) {
            if (nall > maxvatom) { memory->destroy(vatom); memory->create(vatom, nall, 6, "pair:vatom"); maxvatom = nall; }
            for (int i = 0; i < nall; ++i)
                for (int k = 0; k < 6; ++k) vatom[i][k] = 0.0;
        }
    }
};

}  // namespace LAMMPS_NS
