// C ABI of the MI355X MTP evaluator (include/mtp_hip.h).
//
// One context = one potential, one HIP stream and device buffers sized at
// open time for max_atoms x max_nbr (the reference sized its NumPy staging
// pool the same way, zero_overhead.cpp:123-150).  No Python anywhere.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mtp_hip.h"
#include "mtp_format.hpp"
#include "mtp_head.hpp"
#include "mtp_kernels.hpp"
#include "mtp_pack.hpp"
#include "mtp_plan.hpp"
#include "mtp_train.hpp"
#include "nbr_pack.hpp"

using namespace mtp;

static_assert(MTP_NUM_VARIANTS == kNumVariants, "mtp_hip.h must size ms_per_variant for every kernel variant");

namespace {

std::mutex g_err_mu;
std::string g_err;

void set_global_error(const std::string& s)
{
    std::lock_guard<std::mutex> lk(g_err_mu);
    g_err = s;
}

template <typename T>
T* dalloc(size_t n, hipError_t& e)
{
    void* p = nullptr;
    if (e == hipSuccess && n) e = hipMalloc(&p, n * sizeof(T));
    return (T*)p;
}

}  // namespace

// Replay of a call's kernel sequence as a HIP graph.  The key is every
// argument that shapes the sequence (array pointers, sizes, scalars, the
// kernel variant).  A call whose key equals the previous call's is
// captured (the first repeat) and then replayed; any other call runs
// directly and becomes the new reference.  Calls made while the caller's
// stream is being captured, and profiled calls, run directly.
struct GraphCache {
    std::vector<unsigned char> key;
    hipGraphExec_t exec = nullptr;
    void drop()
    {
        if (exec) hipGraphExecDestroy(exec);
        exec = nullptr;
    }
    ~GraphCache() { drop(); }
};

template <class Issue>
int run_cached(GraphCache& gc, const void* key, size_t key_len, hipStream_t s, bool direct, Issue issue)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (direct || hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return issue();
    const unsigned char* k = static_cast<const unsigned char*>(key);
    const bool same = gc.key.size() == key_len && std::memcmp(gc.key.data(), k, key_len) == 0;
    if (!same) {
        gc.drop();
        gc.key.assign(k, k + key_len);
        return issue();
    }
    if (!gc.exec) {
        if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) return issue();
        const int rc = issue();
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(s, &g);
        if (rc != MTP_OK || ec != hipSuccess || !g) {
            if (g) hipGraphDestroy(g);
            gc.key.clear();
            return rc != MTP_OK ? rc : issue();
        }
        const hipError_t ei = hipGraphInstantiate(&gc.exec, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (ei != hipSuccess) {
            gc.exec = nullptr;
            gc.key.clear();
            return issue();
        }
    }
    return hipGraphLaunch(gc.exec, s) == hipSuccess ? MTP_OK : MTP_ERR_HIP;
}

struct mtp_ctx {
    HostPlan plan;
    DevPlan dev{};                        // variant 0 (and the base of the others)
    DevPlan vdev[kNumVariants]{};         // per-variant device plans
    size_t vlds[kNumVariants] = {};          // LDS bytes per variant, 0 = unavailable
    size_t vlds_xi[kNumVariants] = {};       // the same with per-row coefficient vectors
    DevPlan vdev_cache[kNumVariants]{};   // per-row pass over cached moments (layout without tables)
    size_t vlds_cache[kNumVariants] = {};    // its LDS bytes, 0 = unavailable
    int n_cu = 0;
    double* d_eye = nullptr;              // [S][S] identity (basis Jacobians), allocated on first use
    int variant = kVariantV1;
    int max_atoms = 0, max_nbr = 0, debug_level = 0;
    double cutoff = 0.0;
    int radial_fp32 = 0;
    size_t lds_bytes = 0;
    hipStream_t stream = nullptr;
    int device = 0;
    std::string err;
    std::vector<void*> dptrs;
    // parameter and plan buffers
    double *d_radial = nullptr, *d_species = nullptr, *d_moment = nullptr, *d_xi = nullptr;
    // per-call buffers
    double *d_rij = nullptr, *d_Ei = nullptr, *d_g = nullptr, *d_Wi = nullptr, *d_F = nullptr, *d_out = nullptr;
    int *d_it = nullptr, *d_js = nullptr, *d_jt = nullptr, *d_cnt = nullptr, *d_off = nullptr, *d_rev = nullptr;
    int *d_csr_state = nullptr, *d_js_copy = nullptr;  // reuse of the reverse-neighbor CSR (launch_gather)
    int* d_status = nullptr;              // non-finite slot vectors seen by the atom kernels (cleared by the totals)
    double* d_tot = nullptr;
    // pinned staging
    double* h_rij = nullptr;
    int *h_it = nullptr, *h_js = nullptr, *h_jt = nullptr;
    double* h_F = nullptr;
    double* h_out = nullptr;
    // full-list packing on the device (mtp_compute_full_list): grown on demand
    struct Grow {
        void* d = nullptr;
        void* h = nullptr;  // pinned staging of the upload
        size_t bytes = 0;
    };
    Grow g_x, g_type, g_tag, g_ilist, g_off, g_nbr, g_map;
    unsigned long long* d_pack_cnt = nullptr;
    unsigned long long* h_pack_cnt = nullptr;
    int list_nlocal = -1, list_nall = -1, list_inum = -1, list_max_tag = -1;  // the list now on the device
    long long list_nnz = -1;
    int last_host_natoms = 0;             // rows of the last mtp_compute whose E_i, W_i are in d_Ei, d_Wi
    std::vector<double> h_wi;             // staging of mtp_last_atom_values
    // training buffers (allocated at the first training call)
    size_t train_lds = 0;
    DevPlan tstreams{};                   // contraction streams of the training kernel (stream fields only)
    double *d_gxi = nullptr, *d_gc = nullptr, *d_part = nullptr, *d_xisum = nullptr, *d_tsum = nullptr;
    int lds_max = 0;
    GraphCache graph;                     // replay of mtp_compute_padded_device
    unsigned gen = 0;                     // parameter generation (part of the graph keys)
    // live per-kernel timing (events recorded around each atom-kernel launch)
    bool profiling = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pool;
    size_t ev_used = 0;

    int fail(int code, const std::string& msg)
    {
        err = msg;
        if (debug_level > 0) std::fprintf(stderr, "[mtp_hip] %s\n", msg.c_str());
        return code;
    }
    int hip_fail(hipError_t e, const char* where)
    {
        return fail(MTP_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
    }
    template <typename T>
    T* alloc(size_t n, hipError_t& e)
    {
        T* p = dalloc<T>(n, e);
        if (p) dptrs.push_back(p);
        return p;
    }
    ~mtp_ctx()
    {
        if (stream) hipStreamSynchronize(stream);
        for (void* p : dptrs) hipFree(p);
        if (h_rij) hipHostFree(h_rij);
        if (h_it) hipHostFree(h_it);
        if (h_js) hipHostFree(h_js);
        if (h_jt) hipHostFree(h_jt);
        if (h_F) hipHostFree(h_F);
        if (h_out) hipHostFree(h_out);
        for (Grow* g : {&g_x, &g_type, &g_tag, &g_ilist, &g_off, &g_nbr, &g_map}) {
            if (g->d) hipFree(g->d);
            if (g->h) hipHostFree(g->h);
        }
        if (d_pack_cnt) hipFree(d_pack_cnt);
        if (h_pack_cnt) hipHostFree(h_pack_cnt);
        for (auto& pr : ev_pool) { hipEventDestroy(pr.first); hipEventDestroy(pr.second); }
        if (stream) hipStreamDestroy(stream);
    }
};

namespace {

template <typename T>
hipError_t upload(T* d, const std::vector<T>& h, hipStream_t s)
{
    if (h.empty()) return hipSuccess;
    return hipMemcpyAsync(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s);
}

// Uploads the forward and reverse contraction streams of ip into the
// stream fields of dv (device buffers owned by c).
void upload_streams(mtp_ctx* c, const ItemPlan& ip, DevPlan& dv, hipError_t& e)
{
    auto pack = [&](const StreamPlan& sp, int* cb_beg, int* bo_beg, int* xa_beg, const I4** d_wave,
                    const U4** d_tbl, const I4** d_comb, const int** d_bo, const int** d_xa) {
        for (int q = 0; q < kMaxDepth + 2; ++q) {
            cb_beg[q] = sp.cb_beg[q];
            if (bo_beg) bo_beg[q] = sp.bo_beg[q];
            if (xa_beg) xa_beg[q] = sp.xa_beg[q];
        }
        I4* dw = c->alloc<I4>(std::max<size_t>(1, sp.wave.size()), e);
        U4* dt = c->alloc<U4>(std::max<size_t>(1, sp.tbl.size()), e);
        I4* dc = c->alloc<I4>(std::max<size_t>(1, sp.comb.size()), e);
        int* db = c->alloc<int>(std::max<size_t>(1, sp.bo.size()), e);
        int* dx = c->alloc<int>(std::max<size_t>(1, sp.xa.size()), e);
        if (e == hipSuccess) e = upload(dw, sp.wave, c->stream);
        if (e == hipSuccess) e = upload(dt, sp.tbl, c->stream);
        if (e == hipSuccess) e = upload(dc, sp.comb, c->stream);
        if (e == hipSuccess) e = upload(db, sp.bo, c->stream);
        if (e == hipSuccess) e = upload(dx, sp.xa, c->stream);
        *d_wave = dw; *d_tbl = dt; *d_comb = dc;
        if (d_bo) *d_bo = db;
        if (d_xa) *d_xa = dx;
    };
    pack(ip.fwd, dv.fwd_cb_beg, nullptr, nullptr, &dv.fwd_wave, &dv.fwd_tbl, &dv.fwd_comb, nullptr, nullptr);
    pack(ip.rev, dv.rev_cb_beg, dv.rev_bo_beg, dv.rev_xa_beg, &dv.rev_wave, &dv.rev_tbl, &dv.rev_comb, &dv.rev_bo,
         &dv.rev_xa);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);  // the host tables die with ip
    dv.max_partials = ip.max_partials;
}

int upload_params(mtp_ctx* c)
{
    const MtpFile& f = c->plan.file;
    // kernels on any stream may still read the old parameters
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess) e = upload(c->d_radial, f.radial_coeffs, c->stream);
    if (e == hipSuccess) e = upload(c->d_species, f.species_coeffs, c->stream);
    if (e == hipSuccess) e = upload(c->d_moment, f.moment_coeffs, c->stream);
    if (e == hipSuccess) e = upload(c->d_xi, c->plan.xi_of, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return c->hip_fail(e, "parameter upload");
    c->dev.scaling = f.scaling;
    for (auto& v : c->vdev) v.scaling = f.scaling;
    for (auto& v : c->vdev_cache) v.scaling = f.scaling;
    ++c->gen;  // kernel arguments changed: recorded graphs are stale
    return MTP_OK;
}

int open_common(MtpFile&& file, int max_atoms, int max_nbr, double cutoff, int debug_level, mtp_ctx** out)
{
    if (!out) { set_global_error("mtp_open: out is NULL"); return MTP_ERR_ARG; }
    *out = nullptr;
    if (max_atoms <= 0 || max_nbr <= 1 || cutoff <= 0.0) {
        // pair_jax_mtp.cpp:241-243 and 268-269
        set_global_error("max_atoms, max_neighbors, and cutoff must be positive (max_neighbors > 1)");
        return MTP_ERR_ARG;
    }
    if (max_nbr > kAtomBlock) {
        set_global_error("max_neighbors above " + std::to_string(kAtomBlock) + " is not supported");
        return MTP_ERR_UNSUPPORTED;
    }
    auto* c = new mtp_ctx();
    c->max_atoms = max_atoms; c->max_nbr = max_nbr; c->cutoff = cutoff; c->debug_level = debug_level;
    std::string err;
    if (!build_plan(file, 4, c->plan, err)) {
        set_global_error("potential: " + err);
        delete c;
        return MTP_ERR_FILE;
    }
    const HostPlan& p = c->plan;
    const MtpFile& f = p.file;
    c->lds_bytes = atom_lds_bytes(f.alpha_moments_count, max_nbr, f.radial_funcs_count, p.nu_max, p.n_parts);
    hipError_t e = hipGetDevice(&c->device);
    int lds_max = 0;
    if (e == hipSuccess) e = hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device);
    if (e != hipSuccess) {
        set_global_error(std::string("no usable HIP device: ") + hipGetErrorString(e));
        delete c;
        return MTP_ERR_HIP;
    }
    c->lds_max = lds_max;
    if (hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess || c->n_cu <= 0)
        c->n_cu = 256;
    // Moments and adjoints beyond one workgroup's LDS (level 28): the
    // per-target kernel keeps them in a global slice per resident workgroup.
    bool global_moments = false;
    if ((long long)c->lds_bytes > (long long)lds_max) {
        const size_t lg = atom_lds_bytes(f.alpha_moments_count, max_nbr, f.radial_funcs_count, p.nu_max, p.n_parts,
                                         true);
        if ((long long)lg > (long long)lds_max) {
            set_global_error("potential needs " + std::to_string(lg) + " bytes of LDS per atom, device has " +
                             std::to_string(lds_max));
            delete c;
            return MTP_ERR_UNSUPPORTED;
        }
        c->lds_bytes = lg;
        global_moments = true;
    }
    e = set_atom_kernel_lds(kVariantV1, c->lds_bytes);
    if (e == hipSuccess) e = hipStreamCreate(&c->stream);  // default priority
    // plan buffers
    auto* mono = c->alloc<I4>(p.mono.size(), e);
    auto* mono_dst = c->alloc<I4>(p.mono_dst.size(), e);
    auto* ctask = c->alloc<I4>(p.ctask.size(), e);
    auto* fwd_tgt = c->alloc<int>(p.fwd_tgt.size(), e);
    auto* fwd_beg = c->alloc<int>(p.fwd_beg.size(), e);
    auto* fwd_pair = c->alloc<uint32_t>(p.fwd_pair.size(), e);
    auto* fwd_mult = c->alloc<double>(p.fwd_mult.size(), e);
    auto* rev_x = c->alloc<int>(p.rev_x.size(), e);
    auto* rev_beg = c->alloc<int>(p.rev_beg.size(), e);
    auto* rev_pair = c->alloc<uint32_t>(p.rev_pair.size(), e);
    auto* rev_mult = c->alloc<double>(p.rev_mult.size(), e);
    auto* part_blk = c->alloc<I2>(p.part_blk.size(), e);
    auto* mapping = c->alloc<int>(f.alpha_moment_mapping.size(), e);
    auto* map_x = c->alloc<int>(std::max<size_t>(1, p.map_x.size()), e);
    auto* map_beg = c->alloc<int>(p.map_beg.size(), e);
    auto* map_s = c->alloc<int>(std::max<size_t>(1, p.map_s.size()), e);
    c->d_radial = c->alloc<double>(f.radial_coeffs.size(), e);
    c->d_species = c->alloc<double>(f.species_coeffs.size(), e);
    c->d_moment = c->alloc<double>(f.moment_coeffs.size() ? f.moment_coeffs.size() : 1, e);
    c->d_xi = c->alloc<double>(p.xi_of.size(), e);
    // per-call buffers
    const size_t A = (size_t)max_atoms, K = (size_t)max_nbr;
    c->d_rij = c->alloc<double>(A * K * 3, e);
    c->d_it = c->alloc<int>(A, e);
    c->d_js = c->alloc<int>(A * K, e);
    c->d_jt = c->alloc<int>(A * K, e);
    c->d_Ei = c->alloc<double>(A, e);
    c->d_g = c->alloc<double>(A * K * 3, e);
    c->d_Wi = c->alloc<double>(A * 9, e);
    c->d_F = c->alloc<double>(A * 3, e);
    c->d_out = c->alloc<double>(16, e);
    c->d_cnt = c->alloc<int>(2 * A + 2 * ((A + 4095) / 4096) + 16, e);
    c->d_tot = c->alloc<double>(kTotalsBlocks * 10, e);
    c->d_off = c->alloc<int>(A + 1, e);
    c->d_rev = c->alloc<int>(A * K, e);
    c->d_csr_state = c->alloc<int>(8, e);
    c->d_js_copy = c->alloc<int>(A * K, e);
    c->d_status = c->alloc<int>(8, e);
    const int m_grid = global_moments ? (int)std::min<size_t>(A, (size_t)4 * (size_t)c->n_cu) : 0;
    const size_t m_stride = global_moments ? atom_global_moment_doubles(f.alpha_moments_count, max_nbr, p.nu_max,
                                                                       p.n_parts)
                                           : 0;
    double* d_mscratch = global_moments ? c->alloc<double>((size_t)m_grid * m_stride, e) : nullptr;
    if (e == hipSuccess) {  // no previous build yet (valid flag 0); status clear
        static const int zeros[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        e = hipMemcpyAsync(c->d_csr_state, zeros, sizeof(zeros), hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(c->d_status, zeros, sizeof(zeros), hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    }
    if (e == hipSuccess) e = hipHostMalloc((void**)&c->h_rij, A * K * 3 * sizeof(double), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void**)&c->h_it, A * sizeof(int), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void**)&c->h_js, A * K * sizeof(int), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void**)&c->h_jt, A * K * sizeof(int), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void**)&c->h_F, A * 3 * sizeof(double), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void**)&c->h_out, 16 * sizeof(double), hipHostMallocDefault);
    if (e == hipSuccess) e = upload(mono, p.mono, c->stream);
    if (e == hipSuccess) e = upload(mono_dst, p.mono_dst, c->stream);
    if (e == hipSuccess) e = upload(ctask, p.ctask, c->stream);
    if (e == hipSuccess) e = upload(fwd_tgt, p.fwd_tgt, c->stream);
    if (e == hipSuccess) e = upload(fwd_beg, p.fwd_beg, c->stream);
    if (e == hipSuccess) e = upload(fwd_pair, p.fwd_pair, c->stream);
    if (e == hipSuccess) e = upload(fwd_mult, p.fwd_mult, c->stream);
    if (e == hipSuccess) e = upload(rev_x, p.rev_x, c->stream);
    if (e == hipSuccess) e = upload(rev_beg, p.rev_beg, c->stream);
    if (e == hipSuccess) e = upload(rev_pair, p.rev_pair, c->stream);
    if (e == hipSuccess) e = upload(rev_mult, p.rev_mult, c->stream);
    if (e == hipSuccess) e = upload(part_blk, p.part_blk, c->stream);
    if (e == hipSuccess) e = upload(mapping, f.alpha_moment_mapping, c->stream);
    if (e == hipSuccess) e = upload(map_x, p.map_x, c->stream);
    if (e == hipSuccess) e = upload(map_beg, p.map_beg, c->stream);
    if (e == hipSuccess) e = upload(map_s, p.map_s, c->stream);
    if (e != hipSuccess) {
        set_global_error(std::string("device allocation/upload failed: ") + hipGetErrorString(e));
        delete c;
        return MTP_ERR_HIP;
    }
    DevPlan& d = c->dev;
    d.Sp = f.species_count; d.R = f.radial_funcs_count; d.N = f.radial_basis_size;
    d.B = f.B(); d.T = f.T(); d.amc = f.alpha_moments_count; d.S = f.S(); d.nu_max = p.nu_max;
    d.n_live = p.n_live;
    {
        std::vector<char> is_target(p.file.alpha_moments_count, 0);
        for (int x : p.fwd_tgt) is_target[x] = 1;
        d.zero_untargeted = 0;
        for (int x = p.file.B(); x < p.file.alpha_moments_count; ++x)
            if (!is_target[x]) d.zero_untargeted = 1;
    }
    d.scaling = f.scaling; d.min_dist = f.min_dist; d.max_dist = f.max_dist;
    d.radial_coeffs = c->d_radial; d.species_coeffs = c->d_species; d.moment_coeffs = c->d_moment;
    d.mapping = mapping; d.xi_of = c->d_xi;
    std::memcpy(d.blk_off, p.blk_off, sizeof(d.blk_off));
    std::memcpy(d.mu_max_of_nu, p.mu_max_of_nu, sizeof(d.mu_max_of_nu));
    d.n_mono = (int)p.mono.size(); d.mono = mono; d.mono_dst = mono_dst;
    d.n_cwaves = p.n_cwaves; d.ctask = ctask;
    d.n_depth = p.n_depth;
    for (int q = 0; q < kMaxDepth + 2; ++q) { d.fwd_layer[q] = p.fwd_layer[q]; d.rev_layer[q] = p.rev_layer[q]; }
    d.fwd_tgt = fwd_tgt; d.fwd_beg = fwd_beg; d.fwd_pair = fwd_pair; d.fwd_mult = fwd_mult;
    d.rev_x = rev_x; d.rev_beg = rev_beg; d.rev_pair = rev_pair; d.rev_mult = rev_mult;
    d.n_parts = p.n_parts;
    for (int q = 0; q < 9; ++q) d.part_beg[q] = p.part_beg[q];
    d.part_blk = part_blk;
    d.k_layout = max_nbr;
    d.lds_tables = 1;
    d.coef_lds = 0;
    d.n_mapped = (int)p.map_x.size(); d.map_x = map_x; d.map_beg = map_beg; d.map_s = map_s;
    d.mscratch = d_mscratch; d.mscratch_grid = m_grid; d.mscratch_stride = (long long)m_stride;

    // --- kernel variants ------------------------------------------------
    c->vdev[kVariantV1] = d;
    c->vlds[kVariantV1] = c->lds_bytes;
    {
        for (int v : {kVariantV2x256, kVariantV2x512, kVariantV2x64, kVariantPair1024, kVariantPair512, kVariantV2x128}) {
            const int nt = variant_threads(v);
            const bool paired = variant_paired(v);
            HostPlan pp;
            const HostPlan* pv = &p;
            DevPlan dv = d;
            // basic moments by monomial columns in the 256- and 512-thread
            // kernels; the narrower ones keep one monomial per thread
            // (measured faster for them at levels 12-20)
            if (nt < 256) dv.n_cwaves = 0;
            if (variant_parts(v) != p.n_parts) {  // another split of the gradient blocks
                std::string ee;
                build_plan(f, variant_parts(v), pp, ee);  // same validated file: cannot fail
                pv = &pp;
                I2* pb = c->alloc<I2>(pp.part_blk.size(), e);
                if (e == hipSuccess) e = upload(pb, pp.part_blk, c->stream);
                if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
                dv.n_parts = pp.n_parts;
                for (int q = 0; q < 9; ++q) dv.part_beg[q] = pp.part_beg[q];
                dv.part_blk = pb;
            }
            // LDS offsets do not depend on the partial-sum count; keep the
            // partials of a layer within what leaves room for two workgroups
            // per CU when the rest of the layout allows it
            // (one workgroup per CU for the paired rows: the partial sums
            // take what the CU's LDS leaves)
            StreamLayout lay;
            int part_budget;
            if (paired) {
                const AtomLdsPair L0 = atom_lds_pair_layout(f.alpha_moments_count, p.n_live, max_nbr,
                                                            f.radial_funcs_count, p.nu_max, dv.n_parts, nt, 0);
                lay.unit = 16;
                lay.m_off = kPairM;
                lay.dm_off = L0.dm;
                lay.part_off = L0.part;
                lay.absolute = true;
                part_budget = std::max(64, ((int)(lds_max / sizeof(double)) - L0.total_doubles) / 2);
            } else {
                const AtomLds2 L0 = atom_lds2_layout(f.alpha_moments_count, p.n_live, f.B(), max_nbr, f.radial_funcs_count,
                                                     p.nu_max, dv.n_parts, nt, 0, 0);
                const int half_cu = (int)(lds_max / 2 / sizeof(double));
                const int full_cu = (int)(lds_max / sizeof(double));
                // (a layout beyond half the CU's LDS runs one workgroup per CU: its partials get the rest)
                part_budget = std::max(64, (L0.total_doubles < half_cu ? half_cu : full_cu) - L0.total_doubles);
                lay.dm_off = L0.dm;
                lay.part_off = L0.part;
            }
            ItemPlan ip;
            std::string ierr;
            if (!build_items(*pv, nt, lay, part_budget, ip, ierr)) {
                // operand byte offsets beyond 16 bits (level 26: 8,238 live moments): 8-byte slot indices
                if (paired) continue;  // variant unavailable
                lay.absolute = true;
                if (!build_items(*pv, nt, lay, part_budget, ip, ierr)) continue;
            }
            dv.abs_slots = lay.absolute && !paired ? 1 : 0;
            upload_streams(c, ip, dv, e);
            dv.k_layout = max_nbr;
            if (!paired && nt <= 256) {
                // kernels of up to 256 threads take several rounds of their
                // radial loops; they stage the row species' radial
                // coefficients in LDS when that costs no workgroup per CU
                const int coef = f.species_count * f.radial_funcs_count * f.radial_basis_size;
                const auto bytes = [&](int cl) {
                    return atom_lds2_bytes(f.alpha_moments_count, p.n_live, f.B(), max_nbr, f.radial_funcs_count,
                                           p.nu_max, dv.n_parts, nt, ip.max_partials, 0, true, cl);
                };
                set_atom_kernel_lds(v, bytes(coef));  // the occupancy query needs the LDS size allowed
                const int occ0 = atom_kernel_blocks_per_cu(v, bytes(0));
                if (occ0 > 0 && atom_kernel_blocks_per_cu(v, bytes(coef)) >= occ0) dv.coef_lds = coef;
            }
            const size_t lds = paired ? sizeof(double) * (size_t)atom_lds_pair_layout(
                                                f.alpha_moments_count, p.n_live, max_nbr, f.radial_funcs_count,
                                                p.nu_max, dv.n_parts, nt, ip.max_partials).total_doubles
                                      : atom_lds2_bytes(f.alpha_moments_count, p.n_live, f.B(), max_nbr,
                                                        f.radial_funcs_count, p.nu_max, dv.n_parts, nt,
                                                        ip.max_partials, 0, true, dv.coef_lds);
            if (e == hipSuccess && (long long)lds <= (long long)lds_max && set_atom_kernel_lds(v, lds) == hipSuccess) {
                c->vdev[v] = dv;
                c->vlds[v] = lds;
                if (paired) continue;  // plain force evaluation only
                const size_t lxi = atom_lds2_bytes(f.alpha_moments_count, p.n_live, f.B(), max_nbr, f.radial_funcs_count,
                                                   p.nu_max, dv.n_parts, nt, ip.max_partials, f.alpha_moments_count,
                                                   true, dv.coef_lds);
                if ((long long)lxi <= (long long)lds_max && set_atom_kernel_lds(v, lxi) == hipSuccess)
                    c->vlds_xi[v] = lxi;
                // the per-row pass over cached moments (kernels of up to 256
                // threads take that branch): no power table, its own streams
                if (c->vlds_xi[v] && nt <= 256) {
                    // (radial coefficients from global memory: staging them
                    // made this pass slower in measurement)
                    const AtomLds2 Lc = atom_lds2_layout(f.alpha_moments_count, p.n_live, f.B(), max_nbr,
                                                         f.radial_funcs_count, p.nu_max, dv.n_parts, nt, 0,
                                                         f.alpha_moments_count, false, 0);
                    StreamLayout layc;
                    layc.dm_off = Lc.dm;
                    layc.part_off = Lc.part;
                    layc.absolute = lay.absolute;
                    const int half_cu = (int)(lds_max / 2 / sizeof(double));
                    const int budc = std::max(64, (Lc.total_doubles < half_cu ? half_cu : 2 * half_cu) - Lc.total_doubles);
                    ItemPlan ipc;
                    if (build_items(*pv, nt, layc, budc, ipc, ierr)) {
                        DevPlan dvc = dv;
                        upload_streams(c, ipc, dvc, e);
                        dvc.lds_tables = 0;
                        dvc.coef_lds = 0;
                        const size_t lc = atom_lds2_bytes(f.alpha_moments_count, p.n_live, f.B(), max_nbr,
                                                          f.radial_funcs_count, p.nu_max, dvc.n_parts, nt,
                                                          ipc.max_partials, f.alpha_moments_count, false,
                                                          dvc.coef_lds);
                        if (e == hipSuccess && (long long)lc <= (long long)lds_max &&
                            set_atom_kernel_lds(v, lc) == hipSuccess) {
                            c->vdev_cache[v] = dvc;
                            c->vlds_cache[v] = lc;
                        }
                    }
                }
            }
        }
        if (e != hipSuccess) {
            set_global_error(std::string("device allocation/upload failed: ") + hipGetErrorString(e));
            delete c;
            return MTP_ERR_HIP;
        }
        // default: one wavefront per atom for small contraction tables (levels <= 14), two up to level 20,
        // else the widest balanced variant that fits; mtp_autotune measures instead, MTP_KERNEL_VARIANT
        // overrides
        c->variant = c->vlds[kVariantV2x512] ? kVariantV2x512 : (c->vlds[kVariantV2x256] ? kVariantV2x256 : kVariantV1);
        if (f.T() <= 8000 && c->vlds[kVariantV2x128]) c->variant = kVariantV2x128;
        if (f.T() <= 400 && c->vlds[kVariantV2x64]) c->variant = kVariantV2x64;
        if (const char* ev = std::getenv("MTP_KERNEL_VARIANT")) {
            const int v = std::atoi(ev);
            if (v >= 0 && v < kNumVariants && c->vlds[v]) c->variant = v;
        }
    }
    int rc = upload_params(c);
    if (rc) {
        set_global_error(c->err);
        delete c;
        return rc;
    }
    *out = c;
    return MTP_OK;
}

}  // namespace

extern "C" {

const char* mtp_global_error(void)
{
    static thread_local std::string copy;
    std::lock_guard<std::mutex> lk(g_err_mu);
    copy = g_err;
    return copy.c_str();
}

const char* mtp_last_error(const mtp_ctx* ctx) { return ctx ? ctx->err.c_str() : mtp_global_error(); }

int mtp_open(const char* mtp_path, int max_atoms, int max_nbr, double cutoff, int debug_level, mtp_ctx** out)
{
    if (!mtp_path) { set_global_error("mtp_open: path is NULL"); return MTP_ERR_ARG; }
    MtpFile f;
    std::string err;
    if (!read_mtp_file(mtp_path, f, err)) { set_global_error(err); return MTP_ERR_FILE; }
    return open_common(std::move(f), max_atoms, max_nbr, cutoff, debug_level, out);
}

int mtp_open_text(const char* text, int max_atoms, int max_nbr, double cutoff, int debug_level, mtp_ctx** out)
{
    if (!text) { set_global_error("mtp_open_text: text is NULL"); return MTP_ERR_ARG; }
    MtpFile f;
    std::string err;
    if (!parse_mtp_text(text, f, err)) { set_global_error("potential text: " + err); return MTP_ERR_FILE; }
    return open_common(std::move(f), max_atoms, max_nbr, cutoff, debug_level, out);
}

void mtp_close(mtp_ctx* ctx) { delete ctx; }

int mtp_set_precision(mtp_ctx* ctx, int radial_fp32)
{
    if (!ctx) return MTP_ERR_ARG;
    ctx->radial_fp32 = radial_fp32 ? 1 : 0;
    return MTP_OK;
}

int mtp_info(const mtp_ctx* ctx, int* level, int* species_count, int* radial_funcs_count, int* radial_basis_size,
             int* alpha_moments_count, int* basic_count, int* times_count, int* scalar_count)
{
    if (!ctx) return MTP_ERR_ARG;
    const MtpFile& f = ctx->plan.file;
    if (level) *level = ctx->plan.level;
    if (species_count) *species_count = f.species_count;
    if (radial_funcs_count) *radial_funcs_count = f.radial_funcs_count;
    if (radial_basis_size) *radial_basis_size = f.radial_basis_size;
    if (alpha_moments_count) *alpha_moments_count = f.alpha_moments_count;
    if (basic_count) *basic_count = f.B();
    if (times_count) *times_count = f.T();
    if (scalar_count) *scalar_count = f.S();
    return MTP_OK;
}

int mtp_set_params(mtp_ctx* ctx, const double* moment_coeffs, const double* species_coeffs,
                   const double* radial_coeffs, double scaling)
{
    if (!ctx) return MTP_ERR_ARG;
    MtpFile& f = ctx->plan.file;
    if (moment_coeffs) std::memcpy(f.moment_coeffs.data(), moment_coeffs, f.moment_coeffs.size() * sizeof(double));
    if (species_coeffs) std::memcpy(f.species_coeffs.data(), species_coeffs, f.species_coeffs.size() * sizeof(double));
    if (radial_coeffs) std::memcpy(f.radial_coeffs.data(), radial_coeffs, f.radial_coeffs.size() * sizeof(double));
    if (std::isfinite(scaling) && scaling > 0) f.scaling = scaling;
    refresh_plan_params(ctx->plan);
    return upload_params(ctx);
}

int mtp_set_params_device(mtp_ctx* ctx, const double* d_moment_coeffs, const double* d_species_coeffs,
                          const double* d_radial_coeffs, void* stream)
{
    if (!ctx) return MTP_ERR_ARG;
    const MtpFile& f = ctx->plan.file;
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    hipError_t e = hipSuccess;
    auto cp = [&](double* dst, const double* src, size_t n) {
        if (e == hipSuccess && src && n) e = hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, s);
    };
    cp(ctx->d_moment, d_moment_coeffs, f.moment_coeffs.size());
    cp(ctx->d_species, d_species_coeffs, f.species_coeffs.size());
    cp(ctx->d_radial, d_radial_coeffs, f.radial_coeffs.size());
    if (e == hipSuccess && d_moment_coeffs) e = launch_xi_of(ctx->dev, ctx->d_xi, s);
    if (e != hipSuccess) return ctx->hip_fail(e, "device parameter update");
    return MTP_OK;
}

int mtp_compute_padded_device(mtp_ctx* ctx, int nrows, int nslots, const int* d_itypes, const int* d_js,
                              const double* d_rijs, const int* d_jtypes, int cell_rank, double volume,
                              int natoms_energy, int natoms_force, double* d_out, double* d_forces, double* d_eatom,
                              void* stream)
{
    if (!ctx) return MTP_ERR_ARG;
    ctx->last_host_natoms = 0;  // d_Ei, d_Wi no longer hold an mtp_compute's rows
    if (nrows < 0 || nrows > ctx->max_atoms)
        return ctx->fail(MTP_ERR_CAPACITY, "rows " + std::to_string(nrows) + " exceed max_atoms " +
                                               std::to_string(ctx->max_atoms));
    if (nslots <= 0 || nslots > ctx->max_nbr)
        return ctx->fail(MTP_ERR_CAPACITY, "slots " + std::to_string(nslots) + " exceed max_neighbors " +
                                               std::to_string(ctx->max_nbr));
    if (nrows > 0 && (!d_itypes || !d_js || !d_rijs || !d_jtypes || !d_forces))
        return ctx->fail(MTP_ERR_ARG, "null device array");
    if (!d_out) return ctx->fail(MTP_ERR_ARG, "null output");
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    AtomArgs A{};
    A.nrows = nrows; A.nslots = nslots; A.rij = d_rijs; A.itypes = d_itypes; A.jtypes = d_jtypes; A.js = d_js;
    A.Ei = d_eatom ? d_eatom : ctx->d_Ei;
    A.g = ctx->d_g; A.Wi = ctx->d_Wi; A.status = ctx->d_status;
    ForceArgs F{};
    // jax_backend.py:414-418 passes natoms_energy positionally into
    // calc_energy_forces_stress's natoms_force slot, which becomes the mask of
    // accumulate_forces (:279-301): forces and their Newton-3 targets are
    // masked by natoms_energy, and natoms_force has no effect (as nneigh_total).
    (void)natoms_force;
    F.n_e = natoms_energy; F.n_f = natoms_energy; F.cell_rank = cell_rank; F.volume = volume;
    F.cnt = ctx->d_cnt; F.off = ctx->d_off; F.rev = ctx->d_rev; F.F = d_forces; F.out = d_out; F.tot = ctx->d_tot;
    F.csr_state = ctx->d_csr_state; F.js_copy = ctx->d_js_copy; F.status = ctx->d_status;
    hipEvent_t eb = nullptr, ee = nullptr;
    if (ctx->profiling) {
        if (ctx->ev_used == ctx->ev_pool.size()) {
            hipEvent_t a, b;
            if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess)
                return ctx->fail(MTP_ERR_HIP, "event creation failed");
            ctx->ev_pool.push_back({a, b});
        }
        eb = ctx->ev_pool[ctx->ev_used].first;
        ee = ctx->ev_pool[ctx->ev_used].second;
        ++ctx->ev_used;
    }
    struct {
        const void* p[7];
        int n[7];
        double volume;
        hipStream_t s;
    } key;
    std::memset(&key, 0, sizeof(key));
    key.p[0] = d_itypes; key.p[1] = d_js; key.p[2] = d_rijs; key.p[3] = d_jtypes; key.p[4] = d_out;
    key.p[5] = d_forces; key.p[6] = d_eatom;
    key.n[0] = nrows; key.n[1] = nslots; key.n[2] = cell_rank; key.n[3] = natoms_energy;
    key.n[5] = ctx->variant * 2 + (ctx->radial_fp32 != 0);
    key.n[6] = (int)ctx->gen;
    key.volume = volume;
    key.s = s;
    return run_cached(ctx->graph, &key, sizeof(key), s, ctx->profiling, [&]() -> int {
        const hipError_t e = launch_forces(ctx->vdev[ctx->variant], A, F, ctx->vlds[ctx->variant],
                                           ctx->radial_fp32 != 0, s, eb, ee, ctx->variant);
        if (e != hipSuccess) return ctx->hip_fail(e, "kernel launch");
        return MTP_OK;
    });
}

int mtp_set_variant(mtp_ctx* ctx, int variant)
{
    if (!ctx) return MTP_ERR_ARG;
    if (variant < 0 || variant >= kNumVariants || !ctx->vlds[variant])
        return ctx->fail(MTP_ERR_UNSUPPORTED, "kernel variant " + std::to_string(variant) + " is not available");
    ctx->variant = variant;
    return MTP_OK;
}

int mtp_get_variant(const mtp_ctx* ctx, int* variant, long long* lds_bytes)
{
    if (!ctx) return MTP_ERR_ARG;
    if (variant) *variant = ctx->variant;
    if (lds_bytes) *lds_bytes = (long long)ctx->vlds[ctx->variant];
    return MTP_OK;
}

int mtp_autotune(mtp_ctx* ctx, int nrows, int nslots, const int* d_itypes, const int* d_js, const double* d_rijs,
                 const int* d_jtypes, int reps, int* best, double* ms_per_variant)
{
    if (!ctx) return MTP_ERR_ARG;
    ctx->last_host_natoms = 0;  // d_Ei, d_Wi no longer hold an mtp_compute's rows
    if (reps < 1) reps = 1;
    if (nrows <= 0 || nrows > ctx->max_atoms || nslots <= 0 || nslots > ctx->max_nbr)
        return ctx->fail(MTP_ERR_CAPACITY, "autotune shape outside the context's capacity");
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        return ctx->fail(MTP_ERR_HIP, "event creation failed");
    AtomArgs A{};
    A.nrows = nrows; A.nslots = nslots; A.rij = d_rijs; A.itypes = d_itypes; A.jtypes = d_jtypes; A.js = d_js;
    A.Ei = ctx->d_Ei; A.g = ctx->d_g; A.Wi = ctx->d_Wi;
    ForceArgs F{};
    F.n_e = nrows; F.n_f = nrows; F.cell_rank = 3; F.volume = 1.0;
    F.cnt = ctx->d_cnt; F.off = ctx->d_off; F.rev = ctx->d_rev; F.F = ctx->d_F; F.out = ctx->d_out; F.tot = ctx->d_tot;
    F.csr_state = ctx->d_csr_state; F.js_copy = ctx->d_js_copy;
    A.status = ctx->d_status; F.status = ctx->d_status;
    int chosen = ctx->variant;
    double best_ms = 1e300;
    for (int v = 0; v < kNumVariants; ++v) {
        double ms_v = -1.0;
        if (ctx->vlds[v]) {
            hipError_t e = launch_forces(ctx->vdev[v], A, F, ctx->vlds[v], ctx->radial_fp32 != 0, ctx->stream,
                                         nullptr, nullptr, v);
            if (e == hipSuccess) e = hipEventRecord(e0, ctx->stream);
            for (int r = 0; r < reps && e == hipSuccess; ++r)
                e = launch_forces(ctx->vdev[v], A, F, ctx->vlds[v], ctx->radial_fp32 != 0, ctx->stream, nullptr,
                                  nullptr, v);
            if (e == hipSuccess) e = hipEventRecord(e1, ctx->stream);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float ms = 0.0f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
            if (e != hipSuccess) {
                hipEventDestroy(e0);
                hipEventDestroy(e1);
                return ctx->hip_fail(e, "autotune");
            }
            ms_v = ms / reps;
            if (ms_v < best_ms) { best_ms = ms_v; chosen = v; }
        }
        if (ms_per_variant) ms_per_variant[v] = ms_v;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    ctx->variant = chosen;
    if (best) *best = chosen;
    return MTP_OK;
}

int mtp_profile(mtp_ctx* ctx, int enable)
{
    if (!ctx) return MTP_ERR_ARG;
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return ctx->hip_fail(e, "profile");
    ctx->profiling = enable != 0;
    ctx->ev_used = 0;
    return MTP_OK;
}

int mtp_profile_read(mtp_ctx* ctx, double* atom_kernel_ms, int* launches)
{
    if (!ctx) return MTP_ERR_ARG;
    double tot = 0.0;
    for (size_t q = 0; q < ctx->ev_used; ++q) {
        hipError_t e = hipEventSynchronize(ctx->ev_pool[q].second);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev_pool[q].first, ctx->ev_pool[q].second);
        if (e != hipSuccess) return ctx->hip_fail(e, "profile read");
        tot += ms;
    }
    if (atom_kernel_ms) *atom_kernel_ms = tot;
    if (launches) *launches = (int)ctx->ev_used;
    ctx->ev_used = 0;
    return MTP_OK;
}

const double* mtp_last_slot_gradients(const mtp_ctx* ctx) { return ctx ? ctx->d_g : nullptr; }

int mtp_last_atom_values(mtp_ctx* ctx, int natoms, double* eatom, double* vatom6)
{
    if (!ctx) return MTP_ERR_ARG;
    if (natoms < 0 || natoms > ctx->last_host_natoms)
        return ctx->fail(MTP_ERR_ARG, "natoms " + std::to_string(natoms) + " exceeds the last mtp_compute's " +
                                          std::to_string(ctx->last_host_natoms));
    if (natoms == 0 || (!eatom && !vatom6)) return MTP_OK;
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess && eatom) e = hipMemcpy(eatom, ctx->d_Ei, natoms * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess && vatom6) {
        ctx->h_wi.resize((size_t)natoms * 9);
        e = hipMemcpy(ctx->h_wi.data(), ctx->d_Wi, (size_t)natoms * 9 * sizeof(double), hipMemcpyDeviceToHost);
        if (e == hipSuccess) {
            for (int i = 0; i < natoms; ++i) {
                const double* W = ctx->h_wi.data() + (size_t)i * 9;  // W[a * 3 + b]
                double* v = vatom6 + (size_t)i * 6;
                v[0] = -W[0];
                v[1] = -W[4];
                v[2] = -W[8];
                v[3] = -0.5 * (W[1] + W[3]);
                v[4] = -0.5 * (W[2] + W[6]);
                v[5] = -0.5 * (W[5] + W[7]);
            }
        }
    }
    if (e != hipSuccess) return ctx->hip_fail(e, "per-atom values");
    return MTP_OK;
}

int mtp_sync(mtp_ctx* ctx, void* stream)
{
    if (!ctx) return MTP_ERR_ARG;
    hipError_t e = hipStreamSynchronize(stream ? (hipStream_t)stream : ctx->stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "stream synchronize");
    return MTP_OK;
}

int mtp_stage_host(int rows, int max_nbr, int natoms, const double* const* rij_rows, const int* types,
                   const int* const* nbr_rows, const int* nbr_counts, const int* const* nbr_type_rows,
                   int* itypes_out, int* js_out, double* rij_out, int* jtypes_out)
{
    if (rows < 0 || max_nbr <= 0 || natoms < 0 || natoms > rows) return MTP_ERR_ARG;
    if (natoms > 0 && (!rij_rows || !types || !nbr_rows || !nbr_counts || !nbr_type_rows)) return MTP_ERR_ARG;
    if (!itypes_out || !js_out || !rij_out || !jtypes_out) return MTP_ERR_ARG;
    for (int i = 0; i < natoms; ++i)
        if (nbr_counts[i] < 0 || nbr_counts[i] > max_nbr) return MTP_ERR_CAPACITY;
    const size_t tot = (size_t)rows * max_nbr;
    for (size_t q = 0; q < tot; ++q) { js_out[q] = natoms; jtypes_out[q] = -1; }
    for (size_t q = 0; q < tot * 3; ++q) rij_out[q] = 2000.0;
    for (int i = 0; i < rows; ++i) itypes_out[i] = -1;
    if (natoms > 0) std::memcpy(itypes_out, types, (size_t)natoms * sizeof(int));
    for (int i = 0; i < natoms; ++i) {
        const int n = nbr_counts[i];
        const size_t base = (size_t)i * max_nbr;
        if (n > 0) {
            std::memcpy(js_out + base, nbr_rows[i], (size_t)n * sizeof(int));
            std::memcpy(jtypes_out + base, nbr_type_rows[i], (size_t)n * sizeof(int));
            std::memcpy(rij_out + base * 3, rij_rows[i], (size_t)n * 3 * sizeof(double));
        }
    }
    return MTP_OK;
}

int mtp_compute(mtp_ctx* ctx, int natoms, int nneigh_total, const double* const* rij_rows, const int* types,
                const int* const* nbr_rows, const int* nbr_counts, const int* const* nbr_type_rows, double volume,
                double* energy, double* forces, double* stress6, double* virial6)
{
    (void)nneigh_total;
    if (!ctx) return MTP_ERR_ARG;
    if (natoms < 0) return ctx->fail(MTP_ERR_ARG, "natoms is negative");
    if (natoms > ctx->max_atoms)
        return ctx->fail(MTP_ERR_CAPACITY, "natoms " + std::to_string(natoms) + " exceeds max_atoms " +
                                               std::to_string(ctx->max_atoms));
    if (!energy || (natoms > 0 && !forces)) return ctx->fail(MTP_ERR_ARG, "null output pointer");
    const int K = ctx->max_nbr;
    int rc = mtp_stage_host(natoms, K, natoms, rij_rows, types, nbr_rows, nbr_counts, nbr_type_rows, ctx->h_it,
                            ctx->h_js, ctx->h_rij, ctx->h_jt);
    if (rc == MTP_ERR_CAPACITY) return ctx->fail(rc, "a neighbor count exceeds max_neighbors");
    if (rc) return ctx->fail(rc, "null input array");
    hipError_t e = hipSuccess;
    const size_t A = (size_t)natoms;
    if (natoms > 0) {
        e = hipMemcpyAsync(ctx->d_it, ctx->h_it, A * sizeof(int), hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(ctx->d_js, ctx->h_js, A * K * sizeof(int), hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(ctx->d_jt, ctx->h_jt, A * K * sizeof(int), hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(ctx->d_rij, ctx->h_rij, A * K * 3 * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
        if (e != hipSuccess) return ctx->hip_fail(e, "upload");
    }
    rc = mtp_compute_padded_device(ctx, natoms, K, ctx->d_it, ctx->d_js, ctx->d_rij, ctx->d_jt, 3, volume, natoms,
                                   natoms, ctx->d_out, ctx->d_F, nullptr, ctx->stream);
    if (rc) return rc;
    if (natoms > 0) e = hipMemcpyAsync(ctx->h_F, ctx->d_F, A * 3 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(ctx->h_out, ctx->d_out, 14 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "evaluation");
    *energy = ctx->h_out[0];
    if (natoms > 0) std::memcpy(forces, ctx->h_F, A * 3 * sizeof(double));
    if (stress6) std::memcpy(stress6, ctx->h_out + 1, 6 * sizeof(double));
    if (virial6) std::memcpy(virial6, ctx->h_out + 7, 6 * sizeof(double));
    if (ctx->h_out[13] != 0.0) {
        const int st = (int)ctx->h_out[13];
        return ctx->fail(MTP_ERR_NONFINITE, (st & 2) ? "non-finite neighbor vector r_ij in the input rows"
                                                     : "non-finite energy or virial");
    }
    ctx->last_host_natoms = natoms;
    return MTP_OK;
}

}  // extern "C"

namespace {
// Device buffer (and, with pinned, a pinned host buffer) of at least bytes, kept across calls.
hipError_t grow(mtp_ctx::Grow& g, size_t bytes, bool pinned)
{
    if (bytes <= g.bytes && g.d) return hipSuccess;
    if (g.d) hipFree(g.d);
    if (g.h) hipHostFree(g.h);
    g.d = g.h = nullptr;
    g.bytes = 0;
    const size_t b = bytes + bytes / 4 + 256;
    hipError_t e = hipMalloc(&g.d, b);
    if (e == hipSuccess && pinned) e = hipHostMalloc(&g.h, b, hipHostMallocDefault);
    if (e == hipSuccess) g.bytes = b;
    return e;
}
}  // namespace

extern "C" {

int mtp_compute_full_list(mtp_ctx* ctx, int nlocal, int nall, const double* x, const int* type, const int* tag,
                          int inum, const int* ilist, const int* numneigh, const int* const* firstneigh,
                          int list_is_new, const double* prd, const int* periodic, double volume, double* energy,
                          double* forces, double* stress6, double* virial6, long long* n_truncated)
{
    if (!ctx) return MTP_ERR_ARG;
    if (nlocal < 0 || nall < nlocal || inum < 0 || inum > nlocal)
        return ctx->fail(MTP_ERR_ARG, "need 0 <= inum <= nlocal <= nall");
    if (nlocal > ctx->max_atoms)
        return ctx->fail(MTP_ERR_CAPACITY, "natoms " + std::to_string(nlocal) + " exceeds max_atoms " +
                                               std::to_string(ctx->max_atoms));
    if (!energy || (nlocal > 0 && !forces) || !prd || !periodic)
        return ctx->fail(MTP_ERR_ARG, "null output or box pointer");
    if (nall > 0 && (!x || !type || !tag)) return ctx->fail(MTP_ERR_ARG, "null atom array");
    if (inum > 0 && (!ilist || !numneigh || !firstneigh)) return ctx->fail(MTP_ERR_ARG, "null neighbor list");
    const int K = ctx->max_nbr;
    hipStream_t s = ctx->stream;
    // the list, types and tags already on the device are reused when the caller says they are unchanged
    // (LAMMPS: no neighbor-list rebuild since the last call) and the sizes agree
    // host: check ilist, size the flattened list, find the largest local tag
    long long nnz = 0;
    for (int ii = 0; ii < inum; ++ii) {
        const int i = ilist[ii];
        if (i < 0 || i >= nlocal) return ctx->fail(MTP_ERR_ARG, "ilist entry outside the local atoms");
        if (numneigh[i] < 0 || (numneigh[i] > 0 && !firstneigh[i])) return ctx->fail(MTP_ERR_ARG, "bad neighbor row");
        nnz += numneigh[i];
    }
    const bool upload_list = list_is_new || ctx->list_nlocal != nlocal || ctx->list_nall != nall ||
                             ctx->list_inum != inum || ctx->list_nnz != nnz;
    ctx->list_nlocal = ctx->list_nall = ctx->list_inum = -1;  // until this call's upload has been issued
    int max_tag = 0;
    if (upload_list)
        for (int i = 0; i < nlocal; ++i) max_tag = tag[i] > max_tag ? tag[i] : max_tag;
    else
        max_tag = ctx->list_max_tag;
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = grow(ctx->g_x, sizeof(double) * 3 * (size_t)(nall > 0 ? nall : 1), true);
    if (e == hipSuccess) e = grow(ctx->g_type, sizeof(int) * (size_t)(nall > 0 ? nall : 1), true);
    if (e == hipSuccess) e = grow(ctx->g_tag, sizeof(int) * (size_t)(nall > 0 ? nall : 1), true);
    if (e == hipSuccess) e = grow(ctx->g_ilist, sizeof(int) * (size_t)(inum + 1), true);
    if (e == hipSuccess) e = grow(ctx->g_off, sizeof(long long) * (size_t)(inum + 1), true);
    if (e == hipSuccess) e = grow(ctx->g_nbr, sizeof(int) * (size_t)(nnz > 0 ? nnz : 1), true);
    if (e == hipSuccess) e = grow(ctx->g_map, sizeof(int) * ((size_t)max_tag + 1), false);
    if (e == hipSuccess && !ctx->d_pack_cnt) {
        e = hipMalloc((void**)&ctx->d_pack_cnt, 4 * sizeof(unsigned long long));
        if (e == hipSuccess) e = hipHostMalloc((void**)&ctx->h_pack_cnt, 4 * sizeof(unsigned long long), hipHostMallocDefault);
    }
    if (e != hipSuccess) return ctx->hip_fail(e, "full-list buffers");
    auto up = [&](mtp_ctx::Grow& g, size_t bytes) {
        if (e == hipSuccess && bytes) e = hipMemcpyAsync(g.d, g.h, bytes, hipMemcpyHostToDevice, s);
    };
    if (nall > 0) std::memcpy(ctx->g_x.h, x, sizeof(double) * 3 * (size_t)nall);
    up(ctx->g_x, sizeof(double) * 3 * (size_t)nall);
    if (upload_list) {
        // host: the flattened list in ilist order, staged in pinned memory (rows copied by several threads)
        long long* off = static_cast<long long*>(ctx->g_off.h);
        int* nbr = static_cast<int*>(ctx->g_nbr.h);
        off[0] = 0;
        for (int ii = 0; ii < inum; ++ii) off[ii + 1] = off[ii] + numneigh[ilist[ii]];
        const int nth = (int)std::max(1L, std::min<long>(8, nnz / 200000));
        std::vector<unsigned> jmax(nth, 0u);  // neighbor indices past nall would read outside x on the device
        auto copy_rows = [&](int t) {
            unsigned m = 0;
            for (int ii = (int)((long long)inum * t / nth); ii < (int)((long long)inum * (t + 1) / nth); ++ii) {
                const int i = ilist[ii];
                const int n = numneigh[i];
                int* dst = nbr + off[ii];
                if (n > 0) std::memcpy(dst, firstneigh[i], sizeof(int) * (size_t)n);
                for (int q = 0; q < n; ++q) m = std::max(m, (unsigned)(dst[q] & 0x1FFFFFFF));
            }
            jmax[t] = m;
        };
        std::vector<std::thread> pool;
        for (int t = 1; t < nth; ++t) pool.emplace_back(copy_rows, t);
        copy_rows(0);
        for (auto& th : pool) th.join();
        if (nnz > 0 && *std::max_element(jmax.begin(), jmax.end()) >= (unsigned)nall)
            return ctx->fail(MTP_ERR_ARG, "a neighbor index is not below nall");
        if (nall > 0) {
            std::memcpy(ctx->g_type.h, type, sizeof(int) * (size_t)nall);
            std::memcpy(ctx->g_tag.h, tag, sizeof(int) * (size_t)nall);
        }
        if (inum > 0) std::memcpy(ctx->g_ilist.h, ilist, sizeof(int) * (size_t)inum);
        up(ctx->g_type, sizeof(int) * (size_t)nall);
        up(ctx->g_tag, sizeof(int) * (size_t)nall);
        up(ctx->g_ilist, sizeof(int) * (size_t)inum);
        up(ctx->g_off, sizeof(long long) * (size_t)(inum + 1));
        up(ctx->g_nbr, sizeof(int) * (size_t)nnz);
    }
    if (e != hipSuccess) return ctx->hip_fail(e, "full-list upload");
    ctx->list_nlocal = nlocal; ctx->list_nall = nall; ctx->list_inum = inum; ctx->list_max_tag = max_tag;
    ctx->list_nnz = nnz;
    // device: packing + staging into the context's row buffers, then the evaluation
    PackArgs pa{};
    pa.nlocal = nlocal; pa.inum = inum; pa.max_nbr = K; pa.max_tag = max_tag; pa.cutoff = ctx->cutoff;
    for (int q = 0; q < 3; ++q) { pa.prd[q] = prd[q]; pa.periodic[q] = periodic[q]; }
    pa.x = static_cast<const double*>(ctx->g_x.d);
    pa.type = static_cast<const int*>(ctx->g_type.d);
    pa.tag = static_cast<const int*>(ctx->g_tag.d);
    pa.ilist = static_cast<const int*>(ctx->g_ilist.d);
    pa.off = static_cast<const long long*>(ctx->g_off.d);
    pa.nbr = static_cast<const int*>(ctx->g_nbr.d);
    pa.map = static_cast<int*>(ctx->g_map.d);
    pa.itypes = ctx->d_it; pa.js = ctx->d_js; pa.jtypes = ctx->d_jt; pa.rij = ctx->d_rij;
    pa.counters = ctx->d_pack_cnt;
    if (nlocal > 0) {
        e = launch_pack(pa, s);
        if (e != hipSuccess) return ctx->hip_fail(e, "packing launch");
    }
    int rc = mtp_compute_padded_device(ctx, nlocal, K, ctx->d_it, ctx->d_js, ctx->d_rij, ctx->d_jt, 3, volume, nlocal,
                                       nlocal, ctx->d_out, ctx->d_F, nullptr, s);
    if (rc) return rc;
    const size_t A = (size_t)nlocal;
    if (nlocal > 0) e = hipMemcpyAsync(ctx->h_F, ctx->d_F, A * 3 * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(ctx->h_out, ctx->d_out, 14 * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && nlocal > 0)
        e = hipMemcpyAsync(ctx->h_pack_cnt, ctx->d_pack_cnt, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return ctx->hip_fail(e, "evaluation");
    *energy = ctx->h_out[0];
    if (nlocal > 0) std::memcpy(forces, ctx->h_F, A * 3 * sizeof(double));
    if (stress6) std::memcpy(stress6, ctx->h_out + 1, 6 * sizeof(double));
    if (virial6) std::memcpy(virial6, ctx->h_out + 7, 6 * sizeof(double));
    if (n_truncated) *n_truncated = nlocal > 0 ? (long long)ctx->h_pack_cnt[1] : 0;
    if (ctx->h_out[13] != 0.0) {
        const int st = (int)ctx->h_out[13];
        return ctx->fail(MTP_ERR_NONFINITE, (st & 2) ? "non-finite neighbor vector r_ij in the input rows"
                                                     : "non-finite energy or virial");
    }
    ctx->last_host_natoms = nlocal;
    return MTP_OK;
}

int mtp_check_file(const char* mtp_path, int max_nbr, int* level, int* species_count, int* alpha_moments_count,
                   int* basic_count, int* times_count, int* scalar_count, long long* lds_bytes)
{
    if (!mtp_path) { set_global_error("mtp_check_file: path is NULL"); return MTP_ERR_ARG; }
    MtpFile f;
    std::string err;
    if (!read_mtp_file(mtp_path, f, err)) { set_global_error(err); return MTP_ERR_FILE; }
    HostPlan p;
    if (!build_plan(f, 4, p, err)) { set_global_error("potential: " + err); return MTP_ERR_FILE; }
    if (level) *level = p.level;
    if (species_count) *species_count = f.species_count;
    if (alpha_moments_count) *alpha_moments_count = f.alpha_moments_count;
    if (basic_count) *basic_count = f.B();
    if (times_count) *times_count = f.T();
    if (scalar_count) *scalar_count = f.S();
    if (lds_bytes)
        *lds_bytes = (long long)atom_lds_bytes(f.alpha_moments_count, max_nbr, f.radial_funcs_count, p.nu_max,
                                               p.n_parts);
    return MTP_OK;
}

long long mtp_pack_full_list(int nlocal, const double* x, const int* type, const int* tag, int inum,
                             const int* ilist, const int* numneigh, const int* const* firstneigh, double cutoff,
                             const double* prd, const int* periodic, int max_atoms, int max_nbr, double* pos,
                             int* nbr, int* ntype, int* counts, int* atom_types, long long* n_truncated)
{
    if (nlocal < 0 || max_nbr <= 0 || max_atoms <= 0) return MTP_ERR_ARG;
    if (nlocal > max_atoms) return MTP_ERR_CAPACITY;
    if (!x || !type || !tag || !ilist || !numneigh || !firstneigh || !prd || !periodic || !pos || !nbr || !ntype ||
        !counts || !atom_types)
        return MTP_ERR_ARG;
    for (int ii = 0; ii < inum; ++ii)
        if (ilist[ii] < 0 || ilist[ii] >= nlocal) return MTP_ERR_ARG;
    // x is [nall][3] contiguous; build row pointers like LAMMPS double**
    int nall = nlocal;
    for (int ii = 0; ii < inum; ++ii) {
        const int i = ilist[ii];
        for (int jj = 0; jj < numneigh[i]; ++jj) {
            const int j = firstneigh[i][jj] & 0x1FFFFFFF;
            if (j + 1 > nall) nall = j + 1;
        }
    }
    std::vector<const double*> xr((size_t)nall);
    for (int q = 0; q < nall; ++q) xr[q] = x + 3 * (size_t)q;
    OrthoMinImage mi{{prd[0], prd[1], prd[2]}, {periodic[0], periodic[1], periodic[2]}};
    std::vector<int> t2d;
    PackOut o{pos, nbr, ntype, counts, atom_types};
    return pack_full_list(nlocal, xr.data(), type, tag, inum, ilist, numneigh, firstneigh, cutoff, max_atoms,
                          max_nbr, mi, o, t2d, n_truncated);
}

int mtp_batch_forward(mtp_ctx* ctx, int nrows, int nslots, int nconf, const int* d_conf_off, const int* d_conf_of,
                      const int* d_itypes, const int* d_js, const double* d_rijs, const int* d_jtypes,
                      const double* d_volumes, double* d_E, double* d_forces, double* d_stress, void* stream)
{
    (void)d_conf_of;
    if (!ctx) return MTP_ERR_ARG;
    ctx->last_host_natoms = 0;  // d_Ei, d_Wi no longer hold an mtp_compute's rows
    if (nconf <= 0 || !d_conf_off || !d_volumes || !d_E || !d_stress) return ctx->fail(MTP_ERR_ARG, "bad batch");
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    // one evaluation over all rows: per-atom energies, gradients, virials, forces
    int rc = mtp_compute_padded_device(ctx, nrows, nslots, d_itypes, d_js, d_rijs, d_jtypes, 3, 1.0, nrows, nrows,
                                       ctx->d_out, d_forces, ctx->d_Ei, s);
    if (rc) return rc;
    hipError_t e = launch_seg_totals(ctx->d_Ei, ctx->d_Wi, d_conf_off, d_volumes, nconf, d_E, d_stress, s);
    if (e != hipSuccess) return ctx->hip_fail(e, "batch totals");
    return MTP_OK;
}

int mtp_batch_param_grad(mtp_ctx* ctx, int nrows, int nslots, int nconf, const int* d_conf_off, const int* d_conf_of,
                         const int* d_itypes, const int* d_js, const double* d_rijs, const int* d_jtypes,
                         const double* d_volumes, const double* d_gE, const double* d_gF, const double* d_gS,
                         double* d_grad, void* stream)
{
    (void)d_conf_off;
    if (!ctx) return MTP_ERR_ARG;
    ctx->last_host_natoms = 0;  // d_Ei, d_Wi no longer hold an mtp_compute's rows
    if (nrows <= 0 || nrows > ctx->max_atoms)
        return ctx->fail(MTP_ERR_CAPACITY, "batch rows " + std::to_string(nrows) + " outside 1..max_atoms");
    if (nslots <= 0 || nslots > ctx->max_nbr) return ctx->fail(MTP_ERR_CAPACITY, "batch slots exceed max_neighbors");
    if (nconf <= 0 || !d_conf_of || !d_itypes || !d_js || !d_rijs || !d_jtypes || !d_volumes || !d_gE || !d_gF ||
        !d_gS || !d_grad)
        return ctx->fail(MTP_ERR_ARG, "null batch array");
    const HostPlan& p = ctx->plan;
    const MtpFile& f = p.file;
    if (!ctx->d_gxi) {
        // contraction streams of the training kernel: 16-byte (value,
        // tangent) slots; the partial sums get what leaves room for three
        // workgroups per CU, else two, else one
        const int nt = kTrainThreads;
        const TrainLds L0 = train_lds_layout(f.alpha_moments_count, p.n_live, f.B(), ctx->max_nbr,
                                             f.radial_funcs_count, f.radial_basis_size, p.nu_max, p.n_parts, nt, 0);
        StreamLayout lay;
        lay.unit = 16;
        lay.m_off = L0.m;
        lay.dm_off = L0.dm;
        lay.part_off = L0.part;
        lay.absolute = true;
        int budget = 64;
        for (int occ = 3; occ >= 1; --occ) {
            const int b = ((int)(ctx->lds_max / sizeof(double)) / occ - L0.total) / 2;
            if (b >= 64) { budget = b; break; }
        }
        ItemPlan ip;
        std::string ierr;
        if (!build_items(p, nt, lay, budget, ip, ierr))
            return ctx->fail(MTP_ERR_UNSUPPORTED, "training contraction streams: " + ierr);
        ctx->train_lds = sizeof(double) * (size_t)train_lds_layout(f.alpha_moments_count, p.n_live, f.B(),
                                                                   ctx->max_nbr, f.radial_funcs_count,
                                                                   f.radial_basis_size, p.nu_max, p.n_parts, nt,
                                                                   ip.max_partials).total;
        if ((long long)ctx->train_lds > (long long)ctx->lds_max)
            return ctx->fail(MTP_ERR_UNSUPPORTED, "the parameter gradient needs " + std::to_string(ctx->train_lds) +
                                                      " bytes of LDS per atom at this level; the device has " +
                                                      std::to_string(ctx->lds_max));
        hipError_t eu = hipSuccess;
        upload_streams(ctx, ip, ctx->tstreams, eu);
        if (eu != hipSuccess) return ctx->hip_fail(eu, "training streams upload");
        hipError_t e = set_train_kernel_lds(ctx->train_lds);
        const size_t A = (size_t)ctx->max_atoms;
        const size_t ncols = 1 + (size_t)f.species_count * f.radial_funcs_count * f.radial_basis_size;
        const size_t nchunk = (A + kColsumChunk - 1) / kColsumChunk;
        const size_t part = nchunk * std::max((size_t)f.S(), (size_t)f.species_count * ncols);
        ctx->d_gxi = ctx->alloc<double>(A * (size_t)std::max(1, f.S()), e);
        ctx->d_gc = ctx->alloc<double>(A * ncols, e);
        ctx->d_part = ctx->alloc<double>(part, e);
        ctx->d_xisum = ctx->alloc<double>((size_t)std::max(1, f.S()), e);
        ctx->d_tsum = ctx->alloc<double>((size_t)f.species_count * ncols, e);
        if (e != hipSuccess) return ctx->hip_fail(e, "training buffers");
    }
    TrainArgs A{};
    A.nrows = nrows; A.nslots = nslots; A.nconf = nconf;
    A.rij = d_rijs; A.itypes = d_itypes; A.jtypes = d_jtypes; A.js = d_js; A.conf_of = d_conf_of;
    A.volumes = d_volumes; A.gE = d_gE; A.gF = d_gF; A.gS = d_gS;
    A.gxi_part = ctx->d_gxi; A.gc_part = ctx->d_gc;
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    DevPlan tp = ctx->dev;  // parameters and tables of the potential, streams of the training kernel
    const DevPlan& ts = ctx->tstreams;
    tp.fwd_wave = ts.fwd_wave; tp.fwd_tbl = ts.fwd_tbl; tp.fwd_comb = ts.fwd_comb;
    tp.rev_wave = ts.rev_wave; tp.rev_tbl = ts.rev_tbl; tp.rev_comb = ts.rev_comb;
    tp.rev_bo = ts.rev_bo; tp.rev_xa = ts.rev_xa;
    for (int q = 0; q < kMaxDepth + 2; ++q) {
        tp.fwd_cb_beg[q] = ts.fwd_cb_beg[q];
        tp.rev_cb_beg[q] = ts.rev_cb_beg[q];
        tp.rev_bo_beg[q] = ts.rev_bo_beg[q];
        tp.rev_xa_beg[q] = ts.rev_xa_beg[q];
    }
    tp.max_partials = ts.max_partials;
    tp.k_layout = ctx->max_nbr;
    hipError_t e = launch_train_grad(tp, A, ctx->train_lds, ctx->d_part, ctx->d_xisum, ctx->d_tsum, d_grad, s);
    if (e != hipSuccess) return ctx->hip_fail(e, "parameter-gradient launch");
    return MTP_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Basis values, per-row coefficient vectors, basis Jacobians and the
// distilled head.  None of these has a counterpart in the reference (it has
// no learned head); they serve the head's inference and training.
// ---------------------------------------------------------------------------
namespace {

int mode_variant(const mtp_ctx* c, bool with_xi)
{
    const size_t* l = with_xi ? c->vlds_xi : c->vlds;
    if (c->variant != kVariantV1 && !variant_paired(c->variant) && l[c->variant]) return c->variant;
    if (l[kVariantV2x256]) return kVariantV2x256;
    if (l[kVariantV2x512]) return kVariantV2x512;
    return -1;
}

int check_rows(mtp_ctx* c, int nrows, int nslots)
{
    if (nrows < 0 || nrows > c->max_atoms)
        return c->fail(MTP_ERR_CAPACITY, "rows " + std::to_string(nrows) + " outside 0..max_atoms " +
                                             std::to_string(c->max_atoms));
    if (nslots <= 0 || nslots > c->max_nbr)
        return c->fail(MTP_ERR_CAPACITY, "slots " + std::to_string(nslots) + " outside 1..max_neighbors " +
                                             std::to_string(c->max_nbr));
    return MTP_OK;
}

AtomArgs row_args(int nrows, int nslots, const int* it, const double* rij, const int* jt)
{
    AtomArgs A{};
    A.nrows = nrows; A.nslots = nslots; A.rij = rij; A.itypes = it; A.jtypes = jt;
    return A;
}

int slot_grad_xi(mtp_ctx* ctx, int nrows, int nslots, const int* d_itypes, const double* d_rijs, const int* d_jtypes,
                 const double* d_xi, int xi_stride, double* d_g, double* d_Wi, hipStream_t s,
                 const double* d_mom_in = nullptr, const int* d_skip_row = nullptr)
{
    const int v = mode_variant(ctx, true);
    if (v < 0) return ctx->fail(MTP_ERR_UNSUPPORTED, "per-row coefficients need a balanced-item kernel variant");
    AtomArgs A = row_args(nrows, nslots, d_itypes, d_rijs, d_jtypes);
    A.xi_atom = d_xi; A.xi_stride = xi_stride; A.g = d_g; A.Wi = d_Wi; A.mom_in = d_mom_in;
    A.skip_row = d_skip_row;
    const bool cached = d_mom_in && ctx->vlds_cache[v];  // the table-less layout needs the cache
    hipError_t e = launch_atoms(cached ? ctx->vdev_cache[v] : ctx->vdev[v], A,
                                cached ? ctx->vlds_cache[v] : ctx->vlds_xi[v], ctx->radial_fp32, s, v, nrows);
    if (e != hipSuccess) return ctx->hip_fail(e, "slot-gradient launch");
    return MTP_OK;
}

}  // namespace

struct mtp_head {
    mtp_ctx* desc = nullptr;
    mtp_ctx* exact = nullptr;
    HeadNet net{};
    std::string err;
    std::vector<void*> dptrs;
    double *d_theta = nullptr, *d_wt = nullptr, *d_basis = nullptr, *d_xi = nullptr, *d_E = nullptr;
    double* d_mom = nullptr;  // [max_atoms][n_live] moments of the descriptor pass, reread by the adjoint pass
    GraphCache graph;         // replay of mtp_head_compute_padded_device
    int* d_exact_mask = nullptr;      // [max_atoms] rows re-evaluated exactly (skipped by the adjoint pass)
    hipStream_t side = nullptr;       // the exact re-evaluation runs here, beside the adjoint pass
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    int *d_rows = nullptr, *d_count = nullptr;
    double* d_part = nullptr;

    int fail(int code, const std::string& msg)
    {
        err = msg;
        if (desc && desc->debug_level > 0) std::fprintf(stderr, "[mtp_hip] %s\n", msg.c_str());
        return code;
    }
    template <typename T>
    T* alloc(size_t n, hipError_t& e)
    {
        T* p = dalloc<T>(n, e);
        if (p) dptrs.push_back(p);
        return p;
    }
    ~mtp_head()
    {
        if (desc && desc->stream) hipStreamSynchronize(desc->stream);
        if (side) hipStreamSynchronize(side);
        graph.drop();
        if (ev_fork) hipEventDestroy(ev_fork);
        if (ev_join) hipEventDestroy(ev_join);
        if (side) hipStreamDestroy(side);
        for (void* p : dptrs) hipFree(p);
    }
};

extern "C" {

int mtp_descriptor_device(mtp_ctx* ctx, int nrows, int nslots, const int* d_itypes, const double* d_rijs,
                          const int* d_jtypes, double* d_basis, void* stream)
{
    if (!ctx) return MTP_ERR_ARG;
    ctx->last_host_natoms = 0;  // d_Ei, d_Wi no longer hold an mtp_compute's rows
    int rc = check_rows(ctx, nrows, nslots);
    if (rc || nrows == 0) return rc;
    if (!d_itypes || !d_rijs || !d_jtypes || !d_basis) return ctx->fail(MTP_ERR_ARG, "null descriptor array");
    const int v = mode_variant(ctx, false);
    if (v < 0) return ctx->fail(MTP_ERR_UNSUPPORTED, "basis values need a balanced-item kernel variant");
    AtomArgs A = row_args(nrows, nslots, d_itypes, d_rijs, d_jtypes);
    A.basis = d_basis; A.fwd_only = 1;
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    hipError_t e = launch_atoms(ctx->vdev[v], A, ctx->vlds[v], ctx->radial_fp32, s, v, nrows);
    if (e != hipSuccess) return ctx->hip_fail(e, "descriptor launch");
    return MTP_OK;
}

int mtp_slot_grad_xi_device(mtp_ctx* ctx, int nrows, int nslots, const int* d_itypes, const double* d_rijs,
                            const int* d_jtypes, const double* d_xi, int xi_stride, double* d_g, double* d_Wi,
                            void* stream)
{
    if (!ctx) return MTP_ERR_ARG;
    ctx->last_host_natoms = 0;  // d_Ei, d_Wi no longer hold an mtp_compute's rows
    int rc = check_rows(ctx, nrows, nslots);
    if (rc || nrows == 0) return rc;
    if (!d_itypes || !d_rijs || !d_jtypes || !d_xi || !d_g || xi_stride < 0)
        return ctx->fail(MTP_ERR_ARG, "null slot-gradient array or negative stride");
    return slot_grad_xi(ctx, nrows, nslots, d_itypes, d_rijs, d_jtypes, d_xi, xi_stride, d_g, d_Wi,
                        stream ? (hipStream_t)stream : ctx->stream);
}

int mtp_basis_jacobian_device(mtp_ctx* ctx, int nrows, int nslots, const int* d_itypes, const double* d_rijs,
                              const int* d_jtypes, double* d_J, void* stream)
{
    if (!ctx) return MTP_ERR_ARG;
    ctx->last_host_natoms = 0;  // d_Ei, d_Wi no longer hold an mtp_compute's rows
    int rc = check_rows(ctx, nrows, nslots);
    if (rc || nrows == 0) return rc;
    if (!d_itypes || !d_rijs || !d_jtypes || !d_J) return ctx->fail(MTP_ERR_ARG, "null Jacobian array");
    const int S = ctx->plan.file.S();
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    hipError_t e = hipSuccess;
    if (!ctx->d_eye) {
        std::vector<double> eye((size_t)S * S, 0.0);
        for (int q = 0; q < S; ++q) eye[(size_t)q * S + q] = 1.0;
        ctx->d_eye = ctx->alloc<double>(eye.size(), e);
        if (e == hipSuccess) e = upload(ctx->d_eye, eye, s);
        if (e != hipSuccess) return ctx->hip_fail(e, "identity upload");
    }
    const size_t row = (size_t)nslots * 3 * sizeof(double);
    for (int q = 0; q < S; ++q) {
        rc = slot_grad_xi(ctx, nrows, nslots, d_itypes, d_rijs, d_jtypes, ctx->d_eye + (size_t)q * S, 0, ctx->d_g,
                          nullptr, s);
        if (rc) return rc;
        e = hipMemcpy2DAsync(d_J + (size_t)q * nslots * 3, row * S, ctx->d_g, row, row, nrows,
                             hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return ctx->hip_fail(e, "Jacobian copy");
    }
    return MTP_OK;
}

int mtp_jacobian_apply_device(int nrows, int nslots, int S, const double* d_J, const double* d_xi, double* d_g,
                              void* stream)
{
    if (nrows < 0 || nslots <= 0 || S <= 0 || (nrows > 0 && (!d_J || !d_xi || !d_g))) return MTP_ERR_ARG;
    return jacobian_apply(nrows, nslots, S, d_J, d_xi, d_g, (hipStream_t)stream) == hipSuccess ? MTP_OK : MTP_ERR_HIP;
}

int mtp_jacobian_tapply_device(int nrows, int nslots, int S, int n_f, const double* d_J, const int* d_js,
                               const double* d_gF, double* d_v, void* stream)
{
    if (nrows < 0 || nslots <= 0 || S <= 0 || (nrows > 0 && (!d_J || !d_js || !d_gF || !d_v))) return MTP_ERR_ARG;
    if ((size_t)nslots * 3 * sizeof(double) > 64 * 1024) return MTP_ERR_UNSUPPORTED;
    return jacobian_tapply(nrows, nslots, S, n_f, d_J, d_js, d_gF, d_v, (hipStream_t)stream) == hipSuccess
               ? MTP_OK
               : MTP_ERR_HIP;
}

int mtp_gather_forces_device(mtp_ctx* ctx, int nrows, int nslots, const int* d_js, const double* d_g, int n_f,
                             double* d_forces, void* stream)
{
    if (!ctx) return MTP_ERR_ARG;
    ctx->last_host_natoms = 0;  // d_Ei, d_Wi no longer hold an mtp_compute's rows
    int rc = check_rows(ctx, nrows, nslots);
    if (rc || nrows == 0) return rc;
    if (!d_js || !d_g || !d_forces) return ctx->fail(MTP_ERR_ARG, "null gather array");
    AtomArgs A{};
    A.nrows = nrows; A.nslots = nslots; A.js = d_js; A.g = const_cast<double*>(d_g);
    ForceArgs F{};
    F.n_f = n_f; F.cnt = ctx->d_cnt; F.off = ctx->d_off; F.rev = ctx->d_rev; F.F = d_forces; F.tot = ctx->d_tot;
    F.csr_state = ctx->d_csr_state; F.js_copy = ctx->d_js_copy;
    hipError_t e = launch_gather(A, F, stream ? (hipStream_t)stream : ctx->stream);
    if (e != hipSuccess) return ctx->hip_fail(e, "force gather");
    return MTP_OK;
}

}  // extern "C"

// Scratch of mtp_gather_forces_noalloc: the reverse-neighbor CSR (counts and
// cursors, scan tiles, offsets, slot lists) of at most kNoAllocMaxRows rows
// and kNoAllocMaxSlots row slots, in one file-scope device array.  Every GPU
// the library runs on holds its own copy (18.4 MB) until the process ends.
// Host code never uses the array's name as an address: the device address
// comes from hipGetSymbolAddress, once per device.
namespace {
constexpr int kNoAllocMaxRows = 131072;
constexpr long long kNoAllocMaxSlots = 4194304;  // rows x slots per row
constexpr int kNoAllocCnt = 2 * kNoAllocMaxRows + 2 * ((kNoAllocMaxRows + 4095) / 4096) + 16;
constexpr long long kNoAllocInts = kNoAllocCnt + (kNoAllocMaxRows + 1) + kNoAllocMaxSlots;
constexpr int kNoAllocMaxDevices = 64;

__device__ __align__(256) int scratch[kNoAllocInts];

std::mutex g_scratch_mu;
int* g_scratch_ptr[kNoAllocMaxDevices] = {};

// Device address of this device's scratch (looked up on the first call per device).
hipError_t noalloc_scratch(int** out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = static_cast<int*>(p);
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}
}  // namespace

extern "C" {

int mtp_gather_forces_noalloc(int nrows, int nslots, const int* d_js, const double* d_g, int n_f, double* d_forces,
                              void* stream)
{
    if (nrows < 0 || nslots <= 0) {
        set_global_error("mtp_gather_forces_noalloc: rows must be >= 0 and slots > 0");
        return MTP_ERR_ARG;
    }
    if (nrows > kNoAllocMaxRows || (long long)nrows * nslots > kNoAllocMaxSlots) {
        set_global_error("mtp_gather_forces_noalloc: " + std::to_string(nrows) + " rows x " + std::to_string(nslots) +
                         " slots exceed the static scratch (at most " + std::to_string(kNoAllocMaxRows) +
                         " rows and " + std::to_string(kNoAllocMaxSlots) + " slots)");
        return MTP_ERR_CAPACITY;
    }
    if (nrows == 0) return MTP_OK;
    if (!d_js || !d_g || !d_forces) {
        set_global_error("mtp_gather_forces_noalloc: null device array");
        return MTP_ERR_ARG;
    }
    int* base = nullptr;
    hipError_t e = noalloc_scratch(&base);
    if (e != hipSuccess) {
        set_global_error(std::string("mtp_gather_forces_noalloc: scratch address: ") + hipGetErrorString(e));
        return MTP_ERR_HIP;
    }
    AtomArgs A{};
    A.nrows = nrows; A.nslots = nslots; A.js = d_js; A.g = const_cast<double*>(d_g);
    ForceArgs F{};
    F.n_f = n_f; F.F = d_forces;
    F.cnt = base;                           // counts and cursors, cleared by the gather's own fill kernel
    F.off = base + kNoAllocCnt;
    F.rev = base + kNoAllocCnt + (kNoAllocMaxRows + 1);
    F.csr_state = nullptr;                  // no reuse across calls: every call builds its CSR
    e = launch_gather(A, F, (hipStream_t)stream);
    if (e != hipSuccess) {
        set_global_error(std::string("mtp_gather_forces_noalloc: ") + hipGetErrorString(e));
        return MTP_ERR_HIP;
    }
    return MTP_OK;
}

const char* mtp_head_last_error(const mtp_head* h) { return h ? h->err.c_str() : mtp_global_error(); }

int mtp_head_theta_size(void) { return kHeadTheta; }

int mtp_head_open(mtp_ctx* desc, mtp_ctx* exact, const double* theta, const double* mean, const double* inv_std,
                  const double* lo, const double* hi, const double* e_shift, double e_scale, int residual,
                  mtp_head** out)
{
    if (!out) { set_global_error("mtp_head_open: out is NULL"); return MTP_ERR_ARG; }
    *out = nullptr;
    if (!desc || !theta || !mean || !inv_std || !e_shift || (!lo) != (!hi)) {
        set_global_error("mtp_head_open: null descriptor potential or head array");
        return MTP_ERR_ARG;
    }
    const MtpFile& f = desc->plan.file;
    const int S = f.S(), Sp = f.species_count;
    if (S + Sp > kHeadIn) {
        set_global_error("head inputs (basis values + species) exceed " + std::to_string(kHeadIn));
        return MTP_ERR_UNSUPPORTED;
    }
    if (mode_variant(desc, true) < 0 || mode_variant(desc, false) < 0) {
        set_global_error("the descriptor potential has no balanced-item kernel variant");
        return MTP_ERR_UNSUPPORTED;
    }
    if (exact && (exact->plan.file.species_count != Sp || mode_variant(exact, false) < 0)) {
        set_global_error("the exact potential must have the descriptor's species and a balanced-item variant");
        return MTP_ERR_ARG;
    }
    auto* h = new mtp_head();
    h->desc = desc; h->exact = exact;
    const size_t A = (size_t)desc->max_atoms;
    hipError_t e = hipSuccess;
    h->d_theta = h->alloc<double>(kHeadTheta, e);
    h->d_wt = h->alloc<double>(kHeadHidden * kHeadIn + kHeadHidden * kHeadHidden, e);
    double* d_mean = h->alloc<double>(S, e);
    double* d_inv = h->alloc<double>(S, e);
    double* d_lo = lo ? h->alloc<double>((size_t)Sp * S, e) : nullptr;
    double* d_hi = hi ? h->alloc<double>((size_t)Sp * S, e) : nullptr;
    double* d_shift = h->alloc<double>(Sp, e);
    h->d_basis = h->alloc<double>(A * S, e);
    h->d_mom = h->alloc<double>(A * (size_t)std::max(1, desc->plan.n_live), e);
    h->d_exact_mask = h->alloc<int>(A, e);
    if (e == hipSuccess && exact) e = hipStreamCreate(&h->side);  // default priority
    if (e == hipSuccess && exact) e = hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming);
    if (e == hipSuccess && exact) e = hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming);
    h->d_xi = h->alloc<double>(A * S, e);
    h->d_E = h->alloc<double>(A, e);
    h->d_rows = h->alloc<int>(A, e);
    h->d_count = h->alloc<int>(1, e);
    hipStream_t s = desc->stream;
    auto up = [&](double* d, const double* src, size_t n) {
        if (e == hipSuccess && d) e = hipMemcpyAsync(d, src, n * sizeof(double), hipMemcpyHostToDevice, s);
    };
    up(h->d_theta, theta, kHeadTheta);
    up(d_mean, mean, S);
    up(d_inv, inv_std, S);
    up(d_lo, lo, (size_t)Sp * S);
    up(d_hi, hi, (size_t)Sp * S);
    up(d_shift, e_shift, Sp);
    HeadNet& n = h->net;
    n.S = S; n.Sp = Sp; n.theta = h->d_theta; n.wt = h->d_wt; n.mean = d_mean; n.inv_std = d_inv;
    n.lo = d_lo; n.hi = d_hi; n.e_shift = d_shift; n.e_scale = e_scale;
    n.lin = residual ? desc->d_moment : nullptr;
    n.lin_sc = residual ? desc->d_species : nullptr;
    if (e == hipSuccess) e = head_prepare(n, h->d_wt, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        set_global_error(std::string("head allocation/upload failed: ") + hipGetErrorString(e));
        delete h;
        return MTP_ERR_HIP;
    }
    *out = h;
    return MTP_OK;
}

void mtp_head_close(mtp_head* h) { delete h; }

int mtp_head_set_theta(mtp_head* h, const double* theta, int on_device, void* stream)
{
    if (!h) return MTP_ERR_ARG;
    if (!theta) return h->fail(MTP_ERR_ARG, "null theta");
    hipStream_t s = stream ? (hipStream_t)stream : h->desc->stream;
    hipError_t e = hipMemcpyAsync(h->d_theta, theta, kHeadTheta * sizeof(double),
                                  on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = head_prepare(h->net, h->d_wt, s);
    if (e == hipSuccess && !on_device) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return h->fail(MTP_ERR_HIP, std::string("theta upload: ") + hipGetErrorString(e));
    return MTP_OK;
}

int mtp_head_eval_device(mtp_head* h, int nrows, const int* d_itypes, const double* d_basis, double* d_E,
                         double* d_xi, int* d_flags, void* stream)
{
    if (!h) return MTP_ERR_ARG;
    if (nrows < 0 || nrows > h->desc->max_atoms) return h->fail(MTP_ERR_CAPACITY, "rows outside 0..max_atoms");
    if (nrows > 0 && (!d_itypes || !d_basis || !d_E || !d_xi)) return h->fail(MTP_ERR_ARG, "null head array");
    hipError_t e = head_eval(h->net, nrows, d_itypes, d_basis, d_E, d_xi, d_flags, nullptr, nullptr,
                             stream ? (hipStream_t)stream : h->desc->stream);
    if (e != hipSuccess) return h->fail(MTP_ERR_HIP, std::string("head launch: ") + hipGetErrorString(e));
    return MTP_OK;
}

int mtp_head_param_grad_device(mtp_head* h, int nrows, const int* d_itypes, const double* d_basis,
                               const double* d_alpha, const double* d_v, double* d_grad, void* stream)
{
    if (!h) return MTP_ERR_ARG;
    if (nrows <= 0 || nrows > h->desc->max_atoms) return h->fail(MTP_ERR_CAPACITY, "rows outside 1..max_atoms");
    if (!d_itypes || !d_basis || !d_alpha || !d_v || !d_grad) return h->fail(MTP_ERR_ARG, "null head array");
    hipError_t e = hipSuccess;
    if (!h->d_part) {
        h->d_part = h->alloc<double>(head_grad_partials(h->desc->max_atoms), e);
        if (e != hipSuccess) return h->fail(MTP_ERR_HIP, std::string("head partials: ") + hipGetErrorString(e));
    }
    e = head_param_grad(h->net, nrows, d_itypes, d_basis, d_alpha, d_v, h->d_part, d_grad,
                        stream ? (hipStream_t)stream : h->desc->stream);
    if (e != hipSuccess) return h->fail(MTP_ERR_HIP, std::string("head gradient launch: ") + hipGetErrorString(e));
    return MTP_OK;
}

int mtp_head_compute_padded_device(mtp_head* h, int nrows, int nslots, const int* d_itypes, const int* d_js,
                                   const double* d_rijs, const int* d_jtypes, int cell_rank, double volume, int n_e,
                                   int n_f, double* d_out, double* d_forces, double* d_eatom, int* d_flags,
                                   int* d_n_fallback, void* stream)
{
    if (!h) return MTP_ERR_ARG;
    mtp_ctx* D = h->desc;
    mtp_ctx* X = h->exact;
    D->last_host_natoms = 0;  // their d_Ei, d_Wi no longer hold an mtp_compute's rows
    if (X) X->last_host_natoms = 0;
    int rc = check_rows(D, nrows, nslots);
    if (rc) return h->fail(rc, D->err);
    if (X && (nrows > X->max_atoms || nslots > X->max_nbr))
        return h->fail(MTP_ERR_CAPACITY, "rows/slots exceed the exact potential's capacity");
    if (!d_out || (nrows > 0 && (!d_itypes || !d_js || !d_rijs || !d_jtypes || !d_forces)))
        return h->fail(MTP_ERR_ARG, "null input or output array");
    if (n_e < 0 || n_f < 0) return h->fail(MTP_ERR_ARG, "negative natoms_energy / natoms_force");
    hipStream_t s = stream ? (hipStream_t)stream : D->stream;
    double* E = d_eatom ? d_eatom : h->d_E;
    hipError_t e = hipSuccess;
    AtomArgs A = row_args(nrows, nslots, d_itypes, d_rijs, d_jtypes);
    A.js = d_js; A.Ei = E; A.g = D->d_g; A.Wi = D->d_Wi;
    ForceArgs F{};
    // forces masked by natoms_energy, as the plain entry (jax_backend.py:414-418)
    (void)n_f;
    F.n_e = n_e; F.n_f = n_e; F.cell_rank = cell_rank; F.volume = volume;
    F.cnt = D->d_cnt; F.off = D->d_off; F.rev = D->d_rev; F.F = d_forces; F.out = d_out; F.tot = D->d_tot;
    F.csr_state = D->d_csr_state; F.js_copy = D->d_js_copy; F.status = D->d_status;
    const int vd = nrows > 0 ? mode_variant(D, false) : -1, vx = nrows > 0 ? mode_variant(D, true) : -1;
    const int vf = X && nrows > 0 ? mode_variant(X, false) : -1;
    struct {
        const void* p[9];
        int n[11];
        double volume;
        hipStream_t s;
    } key;
    std::memset(&key, 0, sizeof(key));
    key.p[0] = d_itypes; key.p[1] = d_js; key.p[2] = d_rijs; key.p[3] = d_jtypes; key.p[4] = d_out;
    key.p[5] = d_forces; key.p[6] = d_eatom; key.p[7] = d_flags; key.p[8] = d_n_fallback;
    key.n[0] = nrows; key.n[1] = nslots; key.n[2] = cell_rank; key.n[3] = n_e;
    key.n[5] = vd; key.n[6] = vx; key.n[7] = vf; key.n[8] = D->radial_fp32;
    key.n[9] = (int)D->gen; key.n[10] = X ? (int)X->gen : 0;
    key.volume = volume;
    key.s = s;
    return run_cached(h->graph, &key, sizeof(key), s, false, [&]() -> int {
        if (nrows > 0) {
            // 1. invariants  2. network: E_i, dE_i/db_i, out-of-box rows
            {
                const int v = mode_variant(D, false);
                AtomArgs Bd = row_args(nrows, nslots, d_itypes, d_rijs, d_jtypes);
                Bd.basis = h->d_basis; Bd.fwd_only = 1; Bd.mom_out = h->d_mom; Bd.status = D->d_status;
                e = launch_atoms(D->vdev[v], Bd, D->vlds[v], D->radial_fp32, s, v, nrows);
                if (e != hipSuccess)
                    return h->fail(MTP_ERR_HIP, std::string("descriptor launch: ") + hipGetErrorString(e));
            }
            e = head_zero_count(h->d_count, s);
            if (e == hipSuccess)
                e = head_eval(h->net, nrows, d_itypes, h->d_basis, E, h->d_xi, d_flags, X ? h->d_rows : nullptr,
                              X ? h->d_count : nullptr, s, X ? h->d_exact_mask : nullptr);
            if (e != hipSuccess) return h->fail(MTP_ERR_HIP, std::string("head launch: ") + hipGetErrorString(e));
            // 3. exact energies and slot gradients of the out-of-box rows, on
            // the side stream (a handful of workgroups) while 4. the
            // descriptor's adjoint gives the slot gradients of every other row
            if (X) {
                e = hipEventRecord(h->ev_fork, s);
                if (e == hipSuccess) e = hipStreamWaitEvent(h->side, h->ev_fork, 0);
                const int v = mode_variant(X, false);
                AtomArgs B = row_args(nrows, nslots, d_itypes, d_rijs, d_jtypes);
                B.Ei = E; B.g = D->d_g; B.Wi = D->d_Wi; B.rows = h->d_rows; B.row_count = h->d_count;
                const int grid = nrows < X->n_cu ? nrows : X->n_cu;
                if (e == hipSuccess) e = launch_atoms(X->vdev[v], B, X->vlds[v], false, h->side, v, grid);
                if (e == hipSuccess) e = hipEventRecord(h->ev_join, h->side);
                if (e != hipSuccess)
                    return h->fail(MTP_ERR_HIP, std::string("fallback launch: ") + hipGetErrorString(e));
            }
            rc = slot_grad_xi(D, nrows, nslots, d_itypes, d_rijs, d_jtypes, h->d_xi, D->plan.file.S(), D->d_g,
                              D->d_Wi, s, h->d_mom, X ? h->d_exact_mask : nullptr);
            if (rc) return h->fail(rc, D->err);
            if (X) {
                e = hipStreamWaitEvent(s, h->ev_join, 0);
                if (e != hipSuccess) return h->fail(MTP_ERR_HIP, std::string("join: ") + hipGetErrorString(e));
            }
            // 5. Newton-3 gather
            e = launch_gather(A, F, s);
            if (e != hipSuccess) return h->fail(MTP_ERR_HIP, std::string("force gather: ") + hipGetErrorString(e));
        }
        e = launch_totals(A, F, s);
        if (e == hipSuccess && d_n_fallback) {
            if (X && nrows > 0)
                e = hipMemcpyAsync(d_n_fallback, h->d_count, sizeof(int), hipMemcpyDeviceToDevice, s);
            else e = head_zero_count(d_n_fallback, s);
        }
        if (e != hipSuccess) return h->fail(MTP_ERR_HIP, std::string("head totals: ") + hipGetErrorString(e));
        return MTP_OK;
    });
}

}  // extern "C"
