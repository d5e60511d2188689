// fdtd_engine.hip -- MI355X (gfx950) FDTD engine behind the C ABI of
// include/wipems_fdtd.h.
//
// One handle = one Cartesian EC-FDTD domain on one GPU.  Fields live in HBM as
// SoA arrays [3][Nx][Ny][Nzp] (z contiguous, padded to Nzp), so a wave of 64
// lanes reads 64 consecutive z values of one component in one coalesced
// stream.  A timestep is a short chain of kernels on the handle's stream:
//
//   ade_pre_volt  (sparse)  Engine_Ext_LorentzMaterial::DoPreVoltageUpdates
//   volt update   (dense)   Engine::UpdateVoltages with UPML Pre/Post fused
//   tfsf volt     (sparse)  Engine_Ext_TFSF::DoPostVoltageUpdates
//   ade_apply_volt(sparse)  Engine_Ext_Dispersive::Apply2Voltages
//   exc volt      (sparse)  Engine_Ext_Excitation::Apply2Voltages
//   ade_pre_curr / curr update / tfsf curr / ade_apply_curr / exc curr
//   step_end      (tiny)    ++numTS, probe integrals, FD accumulation (FP64)
//
// This is the order of Engine::IterateTS (FDTD/engine.cpp:255-274) with the
// extensions sorted by priority (FDTD/engine.cpp:85-96,212-253): pre hooks run
// Excitation, Lorentz, TFSF, UPML; post/apply hooks run UPML, TFSF, Lorentz,
// Excitation.  The timestep counter lives in device memory so the whole step
// can be captured once in a HIP graph and replayed.
//
// Arithmetic follows the reference operation order exactly,
//   V *= vv;  V += vi * (((a - b) - c) + d)         (FDTD/engine.cpp:126-134)
// with contraction disabled (no FMA) and, for FP32, denormals flushed (the
// SSE engines set FTZ|DAZ, tools/denormal.h:19-31), so an FP32 handle
// reproduces the reference FP32 engine bit for bit.
#include "fdtd_internal.h"
#include <memory>
#include <mutex>
#include <unordered_map>
#include <unordered_set>
#include <string>

thread_local std::string g_last_error;

__global__ void k_energy_final(const double* part, int n, double* out) {
    __shared__ double se[256], sh[256];
    double e = 0.0, h = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) { e += part[2 * i]; h += part[2 * i + 1]; }
    se[threadIdx.x] = e;
    sh[threadIdx.x] = h;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) { se[threadIdx.x] += se[threadIdx.x + s]; sh[threadIdx.x] += sh[threadIdx.x + s]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { out[0] = se[0]; out[1] = sh[0]; }
}

// streaming copy, one 16-byte element per lane with nontemporal loads and
// stores: the calibration kernel for HBM byte counters and the
// achievable-bandwidth reference (6.1 TB/s with plain loads and stores, 6.6
// with nontemporal ones on MI355X; a grid-stride loop of the same copy stays
// near 4.4)
__global__ void __launch_bounds__(256) k_copy(const float4* __restrict__ a, float4* __restrict__ b, long long n) {
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i < n) st_stream(b + i, ld_stream(a + i));
}


namespace wipems {

void invalidate_graph(fdtd_engine* h) {
    if (h->gexec) { (void)hipGraphExecDestroy(h->gexec); h->gexec = nullptr; }
    if (h->graph) { (void)hipGraphDestroy(h->graph); h->graph = nullptr; }
    if (h->gexec1) { (void)hipGraphExecDestroy(h->gexec1); h->gexec1 = nullptr; }
    if (h->graph1) { (void)hipGraphDestroy(h->graph1); h->graph1 = nullptr; }
    h->graph_steps = 0;
}

long long cell_off(const fdtd_engine* h, uint32_t x, uint32_t y, uint32_t z) {
    return (long long)x * h->d.sx + (long long)y * h->d.sy + z;
}

// convert a host array of dtype into the handle's arithmetic type
int convert_host(const fdtd_engine* h, int dtype, const void* src, size_t n, std::vector<char>& dst) {
    if (!src) return fail(FDTD_ERR_ARG, "null array");
    if (dtype != FDTD_DTYPE_F32 && dtype != FDTD_DTYPE_F64) return fail(FDTD_ERR_ARG, "bad dtype %d", dtype);
    dst.resize(n * h->acc);
    if (h->acc == 8) {
        double* o = (double*)dst.data();
        if (dtype == FDTD_DTYPE_F64) memcpy(o, src, n * 8);
        else for (size_t i = 0; i < n; ++i) o[i] = (double)((const float*)src)[i];
    } else {
        float* o = (float*)dst.data();
        if (dtype == FDTD_DTYPE_F32) memcpy(o, src, n * 4);
        else for (size_t i = 0; i < n; ++i) o[i] = (float)((const double*)src)[i];
    }
    return FDTD_OK;
}

// ---------------------------------------------------------------------------
// Large host <-> device copies go through two library-owned pinned chunks
// (hipHostMalloc), double-buffered: the runtime never has to pin (lock) the
// caller's pageable memory for them, and a multi-GB copy holds no
// multi-GB temporary.  One staging area per process, used under a lock.
// ---------------------------------------------------------------------------
struct PinnedStage {
    std::mutex mu;
    char* buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    static constexpr size_t chunk = (size_t)32 << 20;
    int ready() {
        for (int k = 0; k < 2; ++k) {
            if (!buf[k]) HIPCHK(hipHostMalloc((void**)&buf[k], chunk, hipHostMallocDefault));
            if (!ev[k]) HIPCHK(hipEventCreateWithFlags(&ev[k], hipEventDisableTiming));
        }
        return FDTD_OK;
    }
};
static PinnedStage g_stage;
static const size_t kDirectCopy = (size_t)1 << 20;   // smaller copies go straight through

int stage_d2h(hipStream_t st, void* dst, const void* src, size_t bytes) {
    if (bytes <= kDirectCopy) {
        HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        return FDTD_OK;
    }
    std::lock_guard<std::mutex> lk(g_stage.mu);
    int rc = g_stage.ready();
    if (rc) return rc;
    const size_t C = PinnedStage::chunk;
    const size_t nchunks = (bytes + C - 1) / C;
    auto issue = [&](size_t k) -> int {
        const size_t off = k * C, n = std::min(C, bytes - off);
        HIPCHK(hipMemcpyAsync(g_stage.buf[k & 1], (const char*)src + off, n, hipMemcpyDeviceToHost, st));
        HIPCHK(hipEventRecord(g_stage.ev[k & 1], st));
        return FDTD_OK;
    };
    if ((rc = issue(0))) return rc;
    for (size_t k = 0; k < nchunks; ++k) {
        if (k + 1 < nchunks && (rc = issue(k + 1))) return rc;
        HIPCHK(hipEventSynchronize(g_stage.ev[k & 1]));
        const size_t off = k * C, n = std::min(C, bytes - off);
        memcpy((char*)dst + off, g_stage.buf[k & 1], n);
    }
    return FDTD_OK;
}

int stage_h2d(hipStream_t st, void* dst, const void* src, size_t bytes) {
    if (bytes <= kDirectCopy) {
        HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
        return FDTD_OK;
    }
    std::lock_guard<std::mutex> lk(g_stage.mu);
    int rc = g_stage.ready();
    if (rc) return rc;
    const size_t C = PinnedStage::chunk;
    bool used[2] = {false, false};
    for (size_t off = 0, k = 0; off < bytes; off += C, ++k) {
        const size_t n = std::min(C, bytes - off);
        if (used[k & 1]) HIPCHK(hipEventSynchronize(g_stage.ev[k & 1]));   // its previous copy has left the chunk
        memcpy(g_stage.buf[k & 1], (const char*)src + off, n);
        HIPCHK(hipMemcpyAsync((char*)dst + off, g_stage.buf[k & 1], n, hipMemcpyHostToDevice, st));
        HIPCHK(hipEventRecord(g_stage.ev[k & 1], st));
        used[k & 1] = true;
    }
    HIPCHK(hipStreamSynchronize(st));
    return FDTD_OK;
}

// dense [3][Nx][Ny][Nz] host array -> padded device array in arithmetic type
int upload_dense3(fdtd_engine* h, int dtype, const void* src, DevBuf& buf) {
    const size_t nd = (size_t)h->N[0] * h->N[1] * h->N[2];
    int rc = buf.alloc((size_t)3 * h->d.sc * h->acc);
    if (rc) return rc;
    HIPCHK(hipMemsetAsync(buf.p, 0, buf.bytes, h->stream));
    const size_t in_elem = dtype == FDTD_DTYPE_F64 ? 8 : 4;
    // per component: convert on the host, stage the dense rows into a device
    // buffer, then a device-to-device strided copy into the padded layout
    std::vector<char> tmp;
    DevBuf dense;
    if ((rc = dense.alloc(nd * h->acc))) return rc;
    for (int n = 0; n < 3; ++n) {
        const char* s = (const char*)src + (size_t)n * nd * in_elem;
        rc = convert_host(h, dtype, s, nd, tmp);
        if (rc) return rc;
        if ((rc = stage_h2d(h->stream, dense.p, tmp.data(), nd * h->acc))) return rc;
        HIPCHK(hipMemcpy2DAsync((char*)buf.p + (size_t)n * h->d.sc * h->acc, (size_t)h->d.sy * h->acc,
                                dense.p, (size_t)h->N[2] * h->acc, (size_t)h->N[2] * h->acc,
                                (size_t)h->N[0] * h->N[1], hipMemcpyDeviceToDevice, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
    }
    return FDTD_OK;
}

template <typename S>
int launch_step_kernels(fdtd_engine* h, bool timed, bool first, bool last);

int capture_graph(fdtd_engine* h, int steps, hipGraph_t* g, hipGraphExec_t* ge);

int finalize_pml(fdtd_engine* h) {
    if (!h->pml_dirty) return FDTD_OK;
    // the vector kernels apply at most two slabs to one vector of z cells
    const int vz = (int)(16 / h->elem);
    for (int i = 0; i < h->pml.count; ++i)
        for (int j = i + 1; j < h->pml.count; ++j)
            for (int k = j + 1; k < h->pml.count; ++k) {
                const PmlSlab* t[3] = {&h->pml.slab[i], &h->pml.slab[j], &h->pml.slab[k]};
                int lo[3] = {0, 0, 0}, hi[3] = {1 << 30, 1 << 30, 1 << 30};
                for (const PmlSlab* b : t) {
                    for (int n = 0; n < 2; ++n) { lo[n] = std::max(lo[n], b->s[n]); hi[n] = std::min(hi[n], b->s[n] + b->n[n]); }
                    lo[2] = std::max(lo[2], b->s[2] / vz);
                    hi[2] = std::min(hi[2], (b->s[2] + b->n[2] - 1) / vz + 1);
                }
                if (lo[0] < hi[0] && lo[1] < hi[1] && lo[2] < hi[2])
                    return fail(FDTD_ERR_ARG, "UPML slabs %d, %d and %d share a vector of %d z cells", i, j, k, vz);
            }
    const size_t a = h->acc;
    int rc = h->pml_coef.alloc((size_t)6 * h->pml_total * a);
    if (rc) return rc;
    rc = h->pml_flux.alloc((size_t)2 * h->pml_total * a);
    if (rc) return rc;
    HIPCHK(hipMemsetAsync(h->pml_flux.p, 0, h->pml_flux.bytes, h->stream));
    for (int s = 0; s < h->pml.count; ++s) {
        const PmlSlab& b = h->pml.slab[s];
        const size_t cnt = (size_t)3 * b.sc;
        for (int k = 0; k < 6; ++k) {
            const char* src = h->pml_host_coef[s].data() + (size_t)k * cnt * a;
            HIPCHK(hipMemcpyAsync((char*)h->pml_coef.p + ((size_t)k * h->pml_total + b.off) * a, src, cnt * a,
                                  hipMemcpyHostToDevice, h->stream));
        }
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    int rc2 = compress_pml(h);
    if (rc2) return rc2;
    find_shell(h);
    h->pml_host_coef.clear();
    h->pml_dirty = false;
    h->pml_final = true;
    return FDTD_OK;
}

void find_shell(fdtd_engine* h) {
    h->shell_ok = find_interior(h->N, h->pml.slab, h->pml.count, h->ibox);
    // slabs reaching the planes the current update skips (k_curr_upml_faces)
    h->pml_upper = false;
    for (int s = 0; s < h->pml.count; ++s)
        for (int n = 0; n < 3; ++n)
            if (h->pml.slab[s].s[n] + h->pml.slab[s].n[n] >= (int)h->N[n]) h->pml_upper = true;
}

// Deduplicate the per-cell UPML coefficient 9-tuples (k, fo, fn for three
// components) of each side bitwise into at most 65536 rows.  They depend on
// the grading depth and the cell material only, so a slab holds few distinct
// tuples.  Leaves h->pml_cmp false (full arrays only) when there are more.
int compress_pml(fdtd_engine* h) {
    h->pml_cmp = false;
    const size_t a = h->acc;
    const size_t key = 9 * a;
    std::vector<unsigned short> idx[2];
    std::vector<char> tab[2];
    for (int side = 0; side < 2; ++side) {
        idx[side].assign((size_t)h->pml_cells, 0);
        std::unordered_map<uint64_t, std::vector<uint32_t>> seen;
        std::vector<char> rows;   // key bytes per row
        char buf[72];
        for (int s = 0; s < h->pml.count; ++s) {
            const PmlSlab& b = h->pml.slab[s];
            const char* all = h->pml_host_coef[s].data();
            const size_t cnt = (size_t)3 * b.sc;
            for (long long c = 0; c < b.sc; ++c) {
                for (int k = 0; k < 3; ++k)
                    for (int n = 0; n < 3; ++n)
                        memcpy(buf + (k * 3 + n) * a, all + (((size_t)(3 * side + k)) * cnt + (size_t)n * b.sc + c) * a, a);
                uint64_t hsh = 1469598103934665603ull;
                for (size_t i = 0; i < key; ++i) hsh = (hsh ^ (unsigned char)buf[i]) * 1099511628211ull;
                auto& bucket = seen[hsh];
                uint32_t found = UINT32_MAX;
                for (uint32_t r : bucket)
                    if (!memcmp(rows.data() + (size_t)r * key, buf, key)) { found = r; break; }
                if (found == UINT32_MAX) {
                    found = (uint32_t)(rows.size() / key);
                    if (found >= 65536) return FDTD_OK;   // too many: keep the full arrays
                    rows.insert(rows.end(), buf, buf + key);
                    bucket.push_back(found);
                }
                idx[side][(size_t)b.ioff + c] = (unsigned short)found;
            }
        }
        // rows of 12: {k0,k1,k2,0, fo0,fo1,fo2,0, fn0,fn1,fn2,0}
        const size_t nrows = rows.size() / key;
        tab[side].assign(nrows * 12 * a, 0);
        for (size_t r = 0; r < nrows; ++r)
            for (int k = 0; k < 3; ++k)
                memcpy(tab[side].data() + (r * 12 + 4 * k) * a, rows.data() + r * key + k * 3 * a, 3 * a);
    }
    int rc;
    if ((rc = h->pml_vidx.alloc(idx[0].size() * 2)) || (rc = h->pml_cidx.alloc(idx[1].size() * 2)) ||
        (rc = h->pml_vtab.alloc(tab[0].size())) || (rc = h->pml_ctab.alloc(tab[1].size())))
        return rc;
    HIPCHK(hipMemcpyAsync(h->pml_vidx.p, idx[0].data(), idx[0].size() * 2, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->pml_cidx.p, idx[1].data(), idx[1].size() * 2, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->pml_vtab.p, tab[0].data(), tab[0].size(), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->pml_ctab.p, tab[1].data(), tab[1].size(), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->pml_cmp = true;
    return FDTD_OK;
}

bool probes_need_realloc(const fdtd_engine* h, uint32_t upto_ts) {
    if (h->probes_dirty) return true;
    for (auto p : h->probes)
        if (p->td_interval > 0 && (long long)(upto_ts / p->td_interval) + 1 > p->td_slots) return true;
    return false;
}

// Rebuilds the device probe table and grows the TD sample buffers to hold
// samples up to upto_ts.  Old buffers are copied and freed on h->stream: a
// caller that launched step-end kernels reading them on another stream (a
// batch) must wait for that stream first.
int finalize_probes(fdtd_engine* h, uint32_t upto_ts) {
    std::vector<ProbeDev> pd;
    std::vector<long long> off;
    std::vector<int> comp;
    std::vector<signed char> sign;
    if (!probes_need_realloc(h, upto_ts)) return FDTD_OK;
    for (auto p : h->probes) {
        ProbeDev q{};
        q.kind = p->kind;
        q.first = (int)off.size();
        q.count = (int)p->off.size();
        q.td_interval = (int)p->td_interval;
        q.fd_interval = (int)p->fd_interval;
        q.n_freq = (int)p->freqs.size();
        q.weight = p->weight;
        q.t_shift = p->kind == FDTD_PROBE_CURRENT ? 0.5 : 0.0;
        if (p->td_interval > 0) {
            long long need = (long long)(upto_ts / p->td_interval) + 1;
            if (need > p->td_slots) {
                long long ns = std::max(need, 2 * p->td_slots);
                DevBuf nb;
                int rc = nb.alloc((size_t)ns * 2 * sizeof(double));
                if (rc) return rc;
                HIPCHK(hipMemsetAsync(nb.p, 0xff, nb.bytes, h->stream));   // NaN fill = not sampled
                if (p->td.p) HIPCHK(hipMemcpyAsync(nb.p, p->td.p, (size_t)p->td_slots * 2 * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
                HIPCHK(hipStreamSynchronize(h->stream));
                std::swap(p->td.p, nb.p);
                std::swap(p->td.bytes, nb.bytes);
                p->td_slots = ns;
            }
        }
        q.td = p->td.as<double>();
        q.td_slots = p->td_slots;
        q.fd = p->fd.as<double>();
        q.freqs = p->dfreq.as<double>();
        pd.push_back(q);
        off.insert(off.end(), p->off.begin(), p->off.end());
        comp.insert(comp.end(), p->comp.begin(), p->comp.end());
        sign.insert(sign.end(), p->sign.begin(), p->sign.end());
    }
    int rc;
    if ((rc = h->probe_dev.alloc(std::max<size_t>(1, pd.size()) * sizeof(ProbeDev)))) return rc;
    if ((rc = h->e_off.alloc(std::max<size_t>(1, off.size()) * sizeof(long long)))) return rc;
    if ((rc = h->e_comp.alloc(std::max<size_t>(1, comp.size()) * sizeof(int)))) return rc;
    if ((rc = h->e_sign.alloc(std::max<size_t>(1, sign.size())))) return rc;
    if (!pd.empty()) HIPCHK(hipMemcpyAsync(h->probe_dev.p, pd.data(), pd.size() * sizeof(ProbeDev), hipMemcpyHostToDevice, h->stream));
    if (!off.empty()) {
        HIPCHK(hipMemcpyAsync(h->e_off.p, off.data(), off.size() * sizeof(long long), hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipMemcpyAsync(h->e_comp.p, comp.data(), comp.size() * sizeof(int), hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipMemcpyAsync(h->e_sign.p, sign.data(), sign.size(), hipMemcpyHostToDevice, h->stream));
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    h->probes_dirty = false;
    invalidate_graph(h);   // kernel arguments changed
    return FDTD_OK;
}

// one launch per round of distinct locations (TfsfDev::rstart)
template <typename S>
void launch_tfsf(fdtd_engine* h, S* F, const unsigned* stp, const TfsfDev& t, Scales<typename StoreTraits<S>::A> sc, int is_curr) {
    for (int r = 0; r < t.rounds; ++r) {
        const int first = t.rstart[r], end = t.rstart[r + 1];
        if (end <= first) continue;
        hipLaunchKernelGGL((k_tfsf<S>), dim3((end - first + 255) / 256), dim3(256), 0, h->stream, F, stp, t, sc, is_curr, first, end);
    }
}

template <typename S>
int launch_step_end(fdtd_engine* h, int advance) {
    typedef typename StoreTraits<S>::A A;
    hipLaunchKernelGGL((k_step_end<S>), dim3(1), dim3(256), 0, h->stream, h->V.as<S>(), h->I.as<S>(), h->step.as<unsigned>(), advance,
                       h->probe_dev.as<ProbeDev>(), (int)h->probes.size(), h->e_off.as<long long>(), h->e_comp.as<int>(),
                       h->e_sign.as<signed char>(), make_scales<A>(h), h->dT);
    return FDTD_OK;
}

// The skipped current cells on the upper faces (k_curr_upml_faces) stay
// zero unless something writes them: a current-side hook or a caller's
// SetCurr / fdtd_set_field on the current field.
bool faces_live(const fdtd_engine* h) {
    return h->curr_written || h->curr_hook_faces;
}

// whether a current-side hook writes a cell of the planes x = Nx-1,
// y = Ny-1 or z = Nz-1 (element offsets incl. component)
bool on_upper_face(const fdtd_engine* h, long long off) {
    off %= h->d.sc;
    const long long x = off / h->d.sx, r = off - x * h->d.sx, y = r / h->d.sy, z = r - y * h->d.sy;
    return x == (long long)h->N[0] - 1 || y == (long long)h->N[1] - 1 || z == (long long)h->N[2] - 1;
}

void find_curr_hook_faces(fdtd_engine* h) {
    bool hit = false;
    for (long long o : h->exc_host_off[1]) hit |= on_upper_face(h, o);
    for (long long o : h->tfsf_host_off[1]) hit |= on_upper_face(h, o);
    for (auto a : h->ade)
        if (a->flags & 2)
            for (long long o : a->host_off) hit |= on_upper_face(h, o);
    h->curr_hook_faces = hit;
}

// Per-location programs of the fused hooks (k_hooks), built when the hook
// lists change.  Not used when a Mur, lumped-RLC or steady-state hook sits
// between them in the reference's order, or beyond MAX_ADE_ORDERS orders.
int build_hooks(fdtd_engine* h) {
    if (!h->hooks_dirty) return FDTD_OK;
    h->hooks_dirty = false;
    find_curr_hook_faces(h);
    h->hooks_fused = false;
    h->hook[0] = HookProg{};
    h->hook[1] = HookProg{};
    invalidate_graph(h);
    if (h->fuse_hooks_mode == 0 || !h->mur.empty() || h->rlc || h->ss || h->ade.size() > MAX_ADE_ORDERS) return FDTD_OK;
    for (auto a : h->ade)
        if ((size_t)3 * a->count >= (1u << 26)) return FDTD_OK;
    for (int w = 0; w < 2; ++w) {
        std::vector<std::pair<long long, unsigned>> items;
        const std::vector<long long>& tf = h->tfsf_host_off[w];
        for (size_t k = 0; k < tf.size(); ++k) items.emplace_back(tf[k], (unsigned)k);
        // locations of TFSF and excitation entries: an ADE component there
        // joins the location's program, every other one is applied per entry
        std::unordered_set<long long> others(tf.begin(), tf.end());
        if (h->has_exc) others.insert(h->exc_host_off[w].begin(), h->exc_host_off[w].end());
        // a location in two ADE orders is applied in order through its program too
        if (h->ade.size() > 1) {
            std::unordered_map<long long, int> seen;
            for (auto a : h->ade)
                if ((a->flags & (w == 0 ? 1u : 2u)) && a->count)
                    for (long long o0 : a->host_off)
                        for (int n = 0; n < 3; ++n)
                            if (++seen[o0 + n * h->d.sc] == 2) others.insert(o0 + n * h->d.sc);
        }
        for (size_t o = 0; o < h->ade.size(); ++o) {
            AdeHost* a = h->ade[o];
            a->shared[w].alloc(0);
            if (!(a->flags & (w == 0 ? 1u : 2u)) || !a->count) continue;
            std::vector<unsigned char> mask(a->count, 0);
            bool any = false;
            for (uint32_t i = 0; i < a->count; ++i)
                for (int n = 0; n < 3; ++n) {
                    const long long off = a->host_off[i] + n * h->d.sc;
                    if (!others.count(off)) continue;
                    mask[i] |= (unsigned char)(1u << n);
                    any = true;
                    items.emplace_back(off, (1u << 30) | ((unsigned)o << 26) | (unsigned)(n * a->count + i));
                }
            if (any) {
                int rc;
                if ((rc = a->shared[w].alloc(a->count))) return rc;
                HIPCHK(hipMemcpyAsync(a->shared[w].p, mask.data(), a->count, hipMemcpyHostToDevice, h->stream));
            }
        }
        if (h->has_exc) {
            const std::vector<long long>& ex = h->exc_host_off[w];
            for (size_t i = 0; i < ex.size(); ++i) items.emplace_back(ex[i], (2u << 30) | (unsigned)i);
        }
        if (items.size() >= (1u << 30)) return FDTD_OK;
        std::stable_sort(items.begin(), items.end(), [](const std::pair<long long, unsigned>& x, const std::pair<long long, unsigned>& y) {
            return x.first < y.first;
        });
        std::vector<long long> off;
        std::vector<int> start;
        std::vector<unsigned> ops;
        for (size_t k = 0; k < items.size(); ++k) {
            if (k == 0 || items[k].first != items[k - 1].first) { off.push_back(items[k].first); start.push_back((int)k); }
            ops.push_back(items[k].second);
        }
        start.push_back((int)items.size());
        const size_t nt = off.size();
        const size_t b_off = nt * 8, b_start = (nt + 1) * 4, b_ops = ops.size() * 4;
        int rc;
        if ((rc = h->hook_buf[w].alloc(b_off + b_start + b_ops + 16))) return rc;
        char* base = (char*)h->hook_buf[w].p;
        if (nt) {
            HIPCHK(hipMemcpyAsync(base, off.data(), b_off, hipMemcpyHostToDevice, h->stream));
            HIPCHK(hipMemcpyAsync(base + b_off + b_start, ops.data(), b_ops, hipMemcpyHostToDevice, h->stream));
        }
        HIPCHK(hipMemcpyAsync(base + b_off, start.data(), b_start, hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        HookProg& hp = h->hook[w];
        hp.ntarget = (int)nt;
        hp.off = (const long long*)base;
        hp.start = (const int*)(base + b_off);
        hp.ops = (const unsigned*)(base + b_off + b_start);
    }
    h->hooks_fused = true;
    return FDTD_OK;
}

VecLaunch vec_launch(const fdtd_engine* h, int vz, bool curr) {
    VecLaunch L = plan_vec_launch(h->N, h->d.sy, h->pml.count, pml_split_in_use(h), h->ibox, vz, curr, shell_zalign_of(h));
    // shell vectors are numbered in 32 bits, at up to 3x the count in
    // half-width vectors: larger shells run as one UPML launch
    if (L.shell && 3 * (L.g.nA0 + L.g.nA1 + L.g.nB + L.g.nC) >= (1LL << 31))
        L = plan_vec_launch(h->N, h->d.sy, h->pml.count, false, h->ibox, vz, curr);
    return L;
}

// Vectors of the UPML shell kernels: half the interior width for 2-byte
// storage (4 cells, 104 VGPRs instead of 179: 4 waves per SIMD instead of 2;
// large dipole FP16 19.2 instead of 21.6 ms per step, graded 257^3 port
// 0.317 instead of 0.326 ms), the full width for FP32 and FP64 (FP32 at 2
// cells measured 1-2% slower on the large dipole and the dispersive case).
template <typename S> struct ShellWidth {
    enum { value = sizeof(S) == 2 ? VecWidth<S>::value / 2 : VecWidth<S>::value };
};

// Grid of a main launch over nx planes of nvec vectors each, with the tile
// order set in r (a 2D grid of tiles x planes, or a banded 1D grid).
static dim3 vec_grid(const fdtd_engine* h, long long nvec, unsigned nx, VecRange& r) {
    const unsigned tiles = (unsigned)((nvec + 255) / 256);
    r.bpp = 0; r.tpb = 0; r.nx = (int)nx; r.xb = 0;
    int tpb = tile_band_of(h, tiles);
    int xb = tile_xcd_of(h, tiles);
    if (!tpb && !xb) return dim3(tiles, nx, 1);
    if (!tpb) tpb = (int)tiles;
    xb = std::min(xb, (tpb + 7) / 8);     // at most one contiguous sub-band per XCD
    const unsigned bpp = xb ? 8u * (unsigned)xb * (((unsigned)tpb + 8u * xb - 1u) / (8u * xb)) : (unsigned)tpb;
    const unsigned long long nb = (tiles + (unsigned)tpb - 1) / (unsigned)tpb;
    if (nb * bpp * nx >= (1ull << 31)) return dim3(tiles, nx, 1);
    r.tpb = tpb; r.xb = xb; r.bpp = (int)bpp;
    return dim3((unsigned)(nb * bpp * nx), 1, 1);
}

template <typename S, bool PML, int CM>
void launch_volt_main(fdtd_engine* h, const VecLaunch& L, CoeffPtrs<typename StoreTraits<S>::A> cf, PmlPtrs<typename StoreTraits<S>::A> pp) {
    constexpr int VZ = VecWidth<S>::value;
    const dim3 b(256, 1, 1);
    constexpr int VS = ShellWidth<S>::value;
    const ShellGrid gs = VS == VZ ? L.g : shell_grid_split(L, (int)h->N[2], VZ, VZ / VS);
    const long long nshell = L.shell ? gs.nA0 + gs.nA1 + gs.nB + gs.nC : 0;
    if (L.nx) {
        VecRange r = L.r;
        const dim3 g = vec_grid(h, L.nq / VZ, L.nx, r);
        if constexpr (PML && CM == CF_FULL && sizeof(S) >= 4)
            hipLaunchKernelGGL((k_volt_vec_upml_b5<S, CM, VZ>), g, b, 0, h->stream, h->V.as<S>(), h->I.as<S>(), cf, h->d, make_scales<typename StoreTraits<S>::A>(h), h->pml, pp, r);
        else if constexpr (PML)
            hipLaunchKernelGGL((k_volt_vec_upml<S, CM, VZ>), g, b, 0, h->stream, h->V.as<S>(), h->I.as<S>(), cf, h->d, make_scales<typename StoreTraits<S>::A>(h), h->pml, pp, r);
        else
            hipLaunchKernelGGL((k_volt_vec<S, CM, VZ>), g, b, 0, h->stream, h->V.as<S>(), h->I.as<S>(), cf, h->d, make_scales<typename StoreTraits<S>::A>(h), h->pml, pp, r);
    }
    if (nshell > 0)
        hipLaunchKernelGGL((k_volt_shell<S, CM, VS>), dim3((unsigned)((nshell + 255) / 256)), b, 0, h->stream, h->V.as<S>(), h->I.as<S>(), cf, h->d,
                           make_scales<typename StoreTraits<S>::A>(h), h->pml, pp, gs);
}

template <typename S, bool PML, int CM>
void launch_curr_main(fdtd_engine* h, const VecLaunch& L, CoeffPtrs<typename StoreTraits<S>::A> cf, PmlPtrs<typename StoreTraits<S>::A> pp) {
    constexpr int VZ = VecWidth<S>::value;
    const dim3 b(256, 1, 1);
    constexpr int VS = ShellWidth<S>::value;
    const ShellGrid gs = VS == VZ ? L.g : shell_grid_split(L, (int)h->N[2], VZ, VZ / VS);
    const long long nshell = L.shell ? gs.nA0 + gs.nA1 + gs.nB + gs.nC : 0;
    if (L.nx) {
        VecRange r = L.r;
        const dim3 g = vec_grid(h, L.nq / VZ, L.nx, r);
        if constexpr (PML && CM == CF_FULL && sizeof(S) >= 4)
            hipLaunchKernelGGL((k_curr_vec_upml_b5<S, CM, VZ>), g, b, 0, h->stream, h->I.as<S>(), h->V.as<S>(), cf, h->d, make_scales<typename StoreTraits<S>::A>(h), h->pml, pp, r);
        else if constexpr (PML)
            hipLaunchKernelGGL((k_curr_vec_upml<S, CM, VZ>), g, b, 0, h->stream, h->I.as<S>(), h->V.as<S>(), cf, h->d, make_scales<typename StoreTraits<S>::A>(h), h->pml, pp, r);
        else
            hipLaunchKernelGGL((k_curr_vec<S, CM, VZ>), g, b, 0, h->stream, h->I.as<S>(), h->V.as<S>(), cf, h->d, make_scales<typename StoreTraits<S>::A>(h), h->pml, pp, r);
    }
    if (nshell > 0)
        hipLaunchKernelGGL((k_curr_shell<S, CM, VS>), dim3((unsigned)((nshell + 255) / 256)), b, 0, h->stream, h->I.as<S>(), h->V.as<S>(), cf, h->d,
                           make_scales<typename StoreTraits<S>::A>(h), h->pml, pp, gs);
}

template <typename S>
void launch_volt_vec(fdtd_engine* h, CoeffPtrs<typename StoreTraits<S>::A> cf, PmlPtrs<typename StoreTraits<S>::A> pp) {
    const VecLaunch L = vec_launch(h, VecWidth<S>::value, false);
    switch (h->cf_mode) {
    case CF_FULL: if (L.pml_main) launch_volt_main<S, true, CF_FULL>(h, L, cf, pp); else launch_volt_main<S, false, CF_FULL>(h, L, cf, pp); break;
    case CF_IDX16: if (L.pml_main) launch_volt_main<S, true, CF_IDX16>(h, L, cf, pp); else launch_volt_main<S, false, CF_IDX16>(h, L, cf, pp); break;
    case CF_FACT16:
        if constexpr (sizeof(typename StoreTraits<S>::A) == 4) {
            if (L.pml_main) launch_volt_main<S, true, CF_FACT16>(h, L, cf, pp); else launch_volt_main<S, false, CF_FACT16>(h, L, cf, pp);
        }
        break;
    default: if (L.pml_main) launch_volt_main<S, true, CF_IDX32>(h, L, cf, pp); else launch_volt_main<S, false, CF_IDX32>(h, L, cf, pp); break;
    }
}

template <typename S>
void launch_curr_vec(fdtd_engine* h, CoeffPtrs<typename StoreTraits<S>::A> cf, PmlPtrs<typename StoreTraits<S>::A> pp) {
    const VecLaunch L = vec_launch(h, VecWidth<S>::value, true);
    switch (h->cf_mode) {
    case CF_FULL: if (L.pml_main) launch_curr_main<S, true, CF_FULL>(h, L, cf, pp); else launch_curr_main<S, false, CF_FULL>(h, L, cf, pp); break;
    case CF_IDX16: if (L.pml_main) launch_curr_main<S, true, CF_IDX16>(h, L, cf, pp); else launch_curr_main<S, false, CF_IDX16>(h, L, cf, pp); break;
    case CF_FACT16:
        if constexpr (sizeof(typename StoreTraits<S>::A) == 4) {
            if (L.pml_main) launch_curr_main<S, true, CF_FACT16>(h, L, cf, pp); else launch_curr_main<S, false, CF_FACT16>(h, L, cf, pp);
        }
        break;
    default: if (L.pml_main) launch_curr_main<S, true, CF_IDX32>(h, L, cf, pp); else launch_curr_main<S, false, CF_IDX32>(h, L, cf, pp); break;
    }
}

// first / last: the step opens / closes a run of steps launched together
// (a captured graph or one eager loop); with fused hooks the voltage ADE
// pre-update of a step is done by the previous step's current hooks, except
// in the first step of a run
template <typename S>
int launch_step_kernels(fdtd_engine* h, bool timed, bool first, bool last) {
    typedef typename StoreTraits<S>::A A;
    Scales<A> sc = make_scales<A>(h);

    CoeffPtrs<A> cfe{}, cfh{};
    if (h->cf_mode == CF_FULL) {
        cfe.a = h->cf_vv.as<A>(); cfe.b = h->cf_vi.as<A>();
        cfh.a = h->cf_ii.as<A>(); cfh.b = h->cf_iv.as<A>();
    } else {
        cfe.a = h->cf_tabE.as<A>(); cfe.idx = h->cf_idx.p;
        cfh.a = h->cf_tabH.as<A>(); cfh.idx = h->cf_idx.p;
        if (h->cf_mode == CF_FACT16) {
            cfe.g = h->cf_gE.as<A>(); cfh.g = h->cf_gH.as<A>();
            cfe.gy = cfh.gy = h->cf_gy; cfe.gz = cfh.gz = h->cf_gz; cfe.gs = cfh.gs = h->cf_gs;
        }
    }
    PmlPtrs<A> pp{};
    if (h->pml.count) {
        A* c = h->pml_coef.as<A>();
        const long long t = h->pml_total;
        pp.vv = c; pp.vvfo = c + t; pp.vvfn = c + 2 * t;
        pp.ii = c + 3 * t; pp.iifo = c + 4 * t; pp.iifn = c + 5 * t;
        pp.vflux = h->pml_flux.as<A>(); pp.cflux = h->pml_flux.as<A>() + t;
        if (pml_rows_in_use(h)) {
            pp.vidx = h->pml_vidx.as<unsigned short>(); pp.vtab = h->pml_vtab.as<A>();
            pp.cidx = h->pml_cidx.as<unsigned short>(); pp.ctab = h->pml_ctab.as<A>();
        }
    }
    const bool pml = h->pml.count > 0;
    const unsigned* stp = h->step.as<unsigned>();

    // ---- voltage half step ----
    auto ade_dev = [&](AdeHost* a) {
        AdeDev ad{};
        const size_t n3 = (size_t)3 * a->count;
        A* cf = a->coef.as<A>();
        A* st = a->state.as<A>();
        ad.count = (int)a->count;
        ad.off = a->off.as<long long>();
        if (cf) {
            ad.v_int = cf; ad.v_ext = cf + n3; ad.v_lor = cf + 2 * n3;
            ad.i_int = cf + 3 * n3; ad.i_ext = cf + 4 * n3; ad.i_lor = cf + 5 * n3;
        }
        for (int w = 0; w < 2; ++w)
            if (a->ctab[w].p) {
                ad.cidx[w] = a->cidx[w].as<unsigned short>();
                ad.ctab[w] = a->ctab[w].as<A>();
            }
        ad.volt_ade = st; ad.volt_lor = st + n3; ad.curr_ade = st + 2 * n3; ad.curr_lor = st + 3 * n3;
        ad.volt_on = (a->flags & 1) != 0; ad.curr_on = (a->flags & 2) != 0;
        ad.volt_lor_on = (a->flags & 4) != 0; ad.curr_lor_on = (a->flags & 8) != 0;
        ad.shared[0] = a->shared[0].as<unsigned char>();
        ad.shared[1] = a->shared[1].as<unsigned char>();
        return ad;
    };
    // Mur ABC phases: pre in ascending face order, post and apply in
    // descending order (the reference's extension order: equal priorities,
    // faces created x-lo .. z-hi, list reversed), so on shared edges the
    // lower face's value is the one that stays
    auto mur_phase = [&](int phase) {
        const int n = (int)h->mur.size();
        for (int q = 0; q < n; ++q) {
            const MurHost* m = h->mur[phase == 0 ? q : n - 1 - q];
            const long long plane = (long long)m->d.na * m->d.nb;
            hipLaunchKernelGGL((k_mur<S>), dim3((unsigned)((plane + 255) / 256)), dim3(256), 0, h->stream, h->V.as<S>(), stp, m->d, h->d, sc, phase);
        }
    };
    const bool fh = h->hooks_fused;
    AdeSet aset{};
    bool volt_ade = false, curr_ade = false;
    int ade_max = 0;
    for (auto a : h->ade) {
        if (aset.n < MAX_ADE_ORDERS) aset.o[aset.n++] = ade_dev(a);
        volt_ade |= (a->flags & 1) && a->count;
        curr_ade |= (a->flags & 2) && a->count;
        ade_max = std::max(ade_max, (int)a->count);
    }
    // one fused hook launch: the post-update hooks of `which`; with
    // pre_self the ADE pre-update of which's next half step for the
    // components no other hook touches (see k_hooks); pre_next: the other
    // field's ADE pre-update, 1 every component, 2 only the shared ones
    // (the others had theirs in the launch after their own apply)
    auto shared_any = [&](int w) {
        for (auto a : h->ade)
            if ((a->flags & (w == 0 ? 1u : 2u)) && a->count && a->shared[w].p) return true;
        return false;
    };
    auto hooks = [&](int which, int pre_next, bool pre_self) {
        const HookProg& hp = h->hook[which];
        if (pre_next == 2 && !shared_any(1 - which)) pre_next = 0;
        const int n = std::max(std::max(hp.ntarget, pre_next ? ade_max : 0), (which == 0 ? volt_ade : curr_ade) ? ade_max : 0);
        if (!n) return;
        S* F = which == 0 ? h->V.as<S>() : h->I.as<S>();
        const S* G = which == 0 ? h->I.as<S>() : h->V.as<S>();
        hipLaunchKernelGGL((k_hooks<S>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, F, G, stp, hp,
                           which == 0 ? h->tfsf_v : h->tfsf_c, which == 0 ? h->exc_v : h->exc_c, aset, sc, h->d.sc, which, pre_next,
                           pre_self ? 1 : 0);
    };
    mur_phase(0);
    if (!fh || first)
        for (auto a : h->ade)
            if ((a->flags & 1) && a->count)
                hipLaunchKernelGGL((k_ade_pre<S>), dim3((a->count + 255) / 256), dim3(256), 0, h->stream, h->V.as<S>(), ade_dev(a), h->d.sc, sc, 0);
    // lumped RLC: pre after the Lorentz pre hook, apply before the Lorentz
    // apply (the extension is created after Lorentz, so it sits before it in
    // the engine's list, FDTD/engine.cpp:84-88)
    if (h->rlc)
        hipLaunchKernelGGL((k_rlc<S>), dim3((h->rlc->d.count + 255) / 256), dim3(256), 0, h->stream, h->V.as<S>(), stp, h->rlc->d, sc, 0);
    const size_t evi = timed ? h->ev_used : 0;
    if (timed) h->ev_used += 4;
    if (timed) HIPCHK(hipEventRecord(h->ev[evi], h->stream));
    launch_volt_vec<S>(h, cfe, pp);
    if (timed) HIPCHK(hipEventRecord(h->ev[evi + 1], h->stream));
    // the current pre-update of this step: every component at the first step
    // of a launch sequence, else the shared ones (the others were done right
    // after the previous step's current apply)
    if (fh) hooks(0, curr_ade ? (first ? 1 : 2) : 0, volt_ade && !last);
    if (!fh && h->tfsf_v.count)
        launch_tfsf<S>(h, h->V.as<S>(), stp, h->tfsf_v, sc, 0);
    mur_phase(1);
    // steady-state detection: first of the apply hooks (priority +2e6)
    if (h->ss) {
        if (h->ss->d.n)
            hipLaunchKernelGGL((k_ss_record<S>), dim3((h->ss->d.n + 255) / 256), dim3(256), 0, h->stream, h->V.as<S>(), stp, h->ss->d, sc);
        if (h->ss_eval_now) {
            hipLaunchKernelGGL((k_energy_partial<S>), dim3(h->e_blocks), dim3(256), 0, h->stream, h->V.as<S>(), h->I.as<S>(), h->d, sc, h->e_part.as<double>());
            hipLaunchKernelGGL(k_ss_eval, dim3(1), dim3(256), 0, h->stream, h->e_part.as<double>(), h->e_blocks, h->ss->d, kEPS0, kMUE0);
        }
    }
    if (h->rlc)
        hipLaunchKernelGGL((k_rlc<S>), dim3((h->rlc->d.count + 255) / 256), dim3(256), 0, h->stream, h->V.as<S>(), stp, h->rlc->d, sc, 1);
    if (!fh)
        for (auto a : h->ade)
            if ((a->flags & 1) && a->count)
                hipLaunchKernelGGL((k_ade_apply<S>), dim3((a->count + 255) / 256), dim3(256), 0, h->stream, h->V.as<S>(), ade_dev(a), h->d.sc, sc, 0);
    mur_phase(2);
    if (!fh && h->has_exc && h->exc_v.count)
        hipLaunchKernelGGL((k_excite<S>), dim3(std::min(1024, (h->exc_v.count + 255) / 256)), dim3(256), 0, h->stream, h->V.as<S>(), stp, h->exc_v, sc, 0);

    // ---- current half step ----
    if (!fh)
        for (auto a : h->ade)
            if ((a->flags & 2) && a->count)
                hipLaunchKernelGGL((k_ade_pre<S>), dim3((a->count + 255) / 256), dim3(256), 0, h->stream, h->I.as<S>(), ade_dev(a), h->d.sc, sc, 1);
    if (timed) HIPCHK(hipEventRecord(h->ev[evi + 2], h->stream));
    launch_curr_vec<S>(h, cfh, pp);
    if (pml && h->pml_upper && faces_live(h)) {
        const long long nf = (long long)h->N[1] * h->N[2] + (long long)(h->N[0] - 1) * h->N[2] + (long long)(h->N[0] - 1) * (h->N[1] - 1);
        hipLaunchKernelGGL((k_curr_upml_faces<S>), dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, h->stream, h->I.as<S>(), h->d, sc, h->pml, pp);
    }
    if (timed) HIPCHK(hipEventRecord(h->ev[evi + 3], h->stream));
    if (fh) {
        // the next step's voltage pre-update: none after the last step of a
        // launch sequence (the next one starts with k_ade_pre), else the
        // shared components (the others ran after this step's voltage apply)
        hooks(1, volt_ade && !last ? 2 : 0, curr_ade && !last);
        return launch_step_end<S>(h, 1);
    }
    if (h->tfsf_c.count)
        launch_tfsf<S>(h, h->I.as<S>(), stp, h->tfsf_c, sc, 1);
    for (auto a : h->ade)
        if ((a->flags & 2) && a->count)
            hipLaunchKernelGGL((k_ade_apply<S>), dim3((a->count + 255) / 256), dim3(256), 0, h->stream, h->I.as<S>(), ade_dev(a), h->d.sc, sc, 1);
    if (h->has_exc && h->exc_c.count)
        hipLaunchKernelGGL((k_excite<S>), dim3(std::min(1024, (h->exc_c.count + 255) / 256)), dim3(256), 0, h->stream, h->I.as<S>(), stp, h->exc_c, sc, 1);

    return launch_step_end<S>(h, 1);
}

int launch_step(fdtd_engine* h, bool timed, bool first, bool last) {
    switch (h->prec) {
    case FDTD_PREC_FP32: return launch_step_kernels<float>(h, timed, first, last);
    case FDTD_PREC_FP64: return launch_step_kernels<double>(h, timed, first, last);
    case FDTD_PREC_FP16: return launch_step_kernels<__half>(h, timed, first, last);
    default: return launch_step_kernels<__hip_bfloat16>(h, timed, first, last);
    }
}

int capture_graph(fdtd_engine* h, int steps, hipGraph_t* g, hipGraphExec_t* ge) {
    HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    for (int s = 0; s < steps; ++s) {
        int rc = launch_step(h, false, s == 0, s == steps - 1);
        if (rc) { hipGraph_t dummy; (void)hipStreamEndCapture(h->stream, &dummy); return rc; }
    }
    HIPCHK(hipStreamEndCapture(h->stream, g));
    HIPCHK(hipGraphInstantiate(ge, *g, nullptr, nullptr, 0));
    return FDTD_OK;
}

int ready_check(fdtd_engine* h) {
    if (h->cf_mode < 0) return fail(FDTD_ERR_STATE, "operator coefficients not set");
    return FDTD_OK;
}

// ---------------------------------------------------------------------------
// Energy reduction without allocation (fdtd_calc_energy_noalloc): the
// per-block partial sums and the two totals live in one file-scope device
// array per GPU instead of buffers the handle allocates.  The reduction runs
// a fixed number of blocks whatever the domain size, so this array bounds
// the block count only.  Host code never passes `scratch` itself to a kernel
// or a copy: it uses the address hipGetSymbolAddress gives for the device
// that is current, fetched once per device when the first handle on it is
// created and kept in g_scratch_ptr.
// ---------------------------------------------------------------------------
#define ENERGY_BLOCKS 2048
#define SCRATCH_DOUBLES (2 * ENERGY_BLOCKS + 2)
#define MAX_SCRATCH_DEVICES 64
static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[MAX_SCRATCH_DEVICES] = {};

// called with device `dev` current
static int init_scratch(int dev) {
    if (dev < 0 || dev >= MAX_SCRATCH_DEVICES) return fail(FDTD_ERR_RANGE, "device %d beyond the %d the scratch table holds", dev, MAX_SCRATCH_DEVICES);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[dev]) return FDTD_OK;
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(FDTD_ERR_HIP, "hipGetSymbolAddress(scratch) gave a null address on device %d", dev);
    g_scratch_ptr[dev] = (double*)p;
    return FDTD_OK;
}

static double* scratch_of(int dev) {
    if (dev < 0 || dev >= MAX_SCRATCH_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[dev];
}

int ensure_energy_buffers(fdtd_engine* h) {
    int rc;
    if (!h->e_part.p) {
        h->e_blocks = ENERGY_BLOCKS;
        if ((rc = h->e_part.alloc((size_t)h->e_blocks * 2 * sizeof(double)))) return rc;
        if ((rc = h->e_out.alloc(2 * sizeof(double)))) return rc;
    }
    return FDTD_OK;
}

template <typename S>
int absmax_impl(fdtd_engine* h, int which, double out[3]) {
    typedef typename StoreTraits<S>::A A;
    const unsigned blocks = 1024;
    DevBuf part;
    int rc = part.alloc((size_t)3 * blocks * sizeof(double));
    if (rc) return rc;
    hipLaunchKernelGGL((k_absmax<S>), dim3(blocks, 3), dim3(256), 0, h->stream, (which == FDTD_VOLT ? h->V : h->I).template as<S>(),
                       h->d.sc, make_scales<A>(h), which, part.as<double>());
    HIPCHK(hipGetLastError());
    std::vector<double> r((size_t)3 * blocks);
    HIPCHK(hipMemcpyAsync(r.data(), part.p, r.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int n = 0; n < 3; ++n) {
        double m = 0.0;
        for (unsigned b = 0; b < blocks; ++b) m = std::max(m, r[(size_t)n * blocks + b]);
        out[n] = m;
    }
    return FDTD_OK;
}

// the same reduction as energy_impl with its partial sums and totals in the
// device's scratch array; allocates nothing
template <typename S>
int energy_noalloc_impl(fdtd_engine* h, double* part, double* out) {
    typedef typename StoreTraits<S>::A A;
    const int blocks = ENERGY_BLOCKS;
    if ((size_t)2 * blocks + 2 > (size_t)SCRATCH_DOUBLES)
        return fail(FDTD_ERR_RANGE, "energy reduction of %d blocks needs %d doubles of scratch, %d available", blocks, 2 * blocks + 2, SCRATCH_DOUBLES);
    hipLaunchKernelGGL((k_energy_partial<S>), dim3(blocks), dim3(256), 0, h->stream, h->V.as<S>(), h->I.as<S>(), h->d, make_scales<A>(h), part);
    hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(256), 0, h->stream, part, blocks, part + 2 * blocks);
    HIPCHK(hipGetLastError());
    double r[2];
    HIPCHK(hipMemcpyAsync(r, part + 2 * blocks, sizeof(r), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    *out = kEPS0 * r[0] + kMUE0 * r[1];
    return FDTD_OK;
}

template <typename S>
int energy_impl(fdtd_engine* h, double* out) {
    typedef typename StoreTraits<S>::A A;
    int rc;
    if ((rc = ensure_energy_buffers(h))) return rc;
    hipLaunchKernelGGL((k_energy_partial<S>), dim3(h->e_blocks), dim3(256), 0, h->stream, h->V.as<S>(), h->I.as<S>(), h->d, make_scales<A>(h), h->e_part.as<double>());
    hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(256), 0, h->stream, h->e_part.as<double>(), h->e_blocks, h->e_out.as<double>());
    double r[2];
    HIPCHK(hipMemcpyAsync(r, h->e_out.p, sizeof(r), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    *out = kEPS0 * r[0] + kMUE0 * r[1];
    return FDTD_OK;
}

// edge lengths and stored materials of the dump kernels
DftAxes dump_axes(const fdtd_engine* h) {
    const int tot = (int)(h->N[0] + h->N[1] + h->N[2]);
    DftAxes ax;
    ax.len = h->edge_len.as<double>();
    ax.dlen = h->edge_len.as<double>() + tot;
    ax.off[0] = 0; ax.off[1] = (int)h->N[0]; ax.off[2] = (int)(h->N[0] + h->N[1]);
    for (int n = 0; n < 3; ++n) ax.N[n] = (int)h->N[n];
    return ax;
}

DumpMat dump_mat(const fdtd_engine* h) {
    DumpMat m;
    m.eps = h->mat[FDTD_MAT_EPS].p ? h->mat[FDTD_MAT_EPS].as<float>() : nullptr;
    m.kappa = h->mat[FDTD_MAT_KAPPA].p ? h->mat[FDTD_MAT_KAPPA].as<float>() : nullptr;
    m.mu = h->mat[FDTD_MAT_MU].p ? h->mat[FDTD_MAT_MU].as<float>() : nullptr;
    return m;
}

template <typename S>
int launch_dft(fdtd_engine* h, DftHost* b, uint32_t ts) {
    typedef typename StoreTraits<S>::A A;
    DftDev dd;
    dd.type = b->type;
    dd.interp = b->interp;
    for (int n = 0; n < 3; ++n) { dd.s[n] = (int)b->s[n]; dd.n[n] = (int)b->n[n]; }
    dd.nodes = b->nodes;
    dd.n_freq = (int)b->freqs.size();
    dd.freqs = b->dfreq.as<double>();
    dd.acc = b->acc.as<double2>();
    // Engine_Interface_FDTD::GetTime (engine_interface_fdtd.h:52); H uses the dual time
    const double t = ((double)ts + (b->type == FDTD_DUMP_H ? 0.5 : 0.0)) * h->dT;
    const double g = h->dT * (double)b->interval;
    const long long blocks = (b->nodes + 255) / 256;
    if (blocks > 0x7fffffffLL) return fail(FDTD_ERR_RANGE, "DFT box too large");
    hipLaunchKernelGGL((k_dft<S>), dim3((unsigned)blocks), dim3(256), 0, h->stream, h->V.as<S>(), h->I.as<S>(), h->d, make_scales<A>(h),
                       dump_axes(h), dump_mat(h), dd, t, g);
    ++b->samples;
    return FDTD_OK;
}

// hands the sample in host slot `slot` of box b to its sink
int drain_td_slot(fdtd_engine* h, TdHost* b, int slot) {
    if (!b->pending[slot]) return FDTD_OK;
    HIPCHK(hipEventSynchronize(b->ev[slot]));
    b->pending[slot] = false;
    const double t = ((double)b->pts[slot] + (b->type == FDTD_DUMP_H ? 0.5 : 0.0)) * h->dT;
    b->sink(b->user, b->id, b->pts[slot], t, b->host[slot], (uint64_t)3 * b->nodes);
    return FDTD_OK;
}

// every pending sample, oldest first per box
int drain_tds(fdtd_engine* h) {
    int rc;
    for (auto b : h->td)
        for (int q = 0; q < 2; ++q)
            if ((rc = drain_td_slot(h, b, (int)((b->count + q) & 1)))) return rc;
    return FDTD_OK;
}

template <typename S>
int launch_td(fdtd_engine* h, TdHost* b, uint32_t ts) {
    typedef typename StoreTraits<S>::A A;
    const int slot = (int)(b->count & 1);
    int rc;
    if ((rc = drain_td_slot(h, b, slot))) return rc;
    DftDev dd{};
    dd.type = b->type;
    dd.interp = b->interp;
    for (int n = 0; n < 3; ++n) { dd.s[n] = (int)b->s[n]; dd.n[n] = (int)b->n[n]; }
    dd.nodes = b->nodes;
    hipLaunchKernelGGL((k_snapshot<S>), dim3((unsigned)((b->nodes + 255) / 256)), dim3(256), 0, h->stream, h->V.as<S>(), h->I.as<S>(), h->d,
                       make_scales<A>(h), dump_axes(h), dump_mat(h), dd, b->buf.as<float>());
    HIPCHK(hipMemcpyAsync(b->host[slot], b->buf.p, (size_t)3 * b->nodes * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipEventRecord(b->ev[slot], h->stream));
    b->pending[slot] = true;
    b->pts[slot] = ts;
    ++b->count;
    return FDTD_OK;
}

int launch_tds(fdtd_engine* h, uint32_t ts) {
    for (auto b : h->td) {
        if (ts % b->interval) continue;
        int rc;
        switch (h->prec) {
        case FDTD_PREC_FP32: rc = launch_td<float>(h, b, ts); break;
        case FDTD_PREC_FP64: rc = launch_td<double>(h, b, ts); break;
        case FDTD_PREC_FP16: rc = launch_td<__half>(h, b, ts); break;
        default: rc = launch_td<__hip_bfloat16>(h, b, ts); break;
        }
        if (rc) return rc;
    }
    HIPCHK(hipGetLastError());
    return FDTD_OK;
}

int launch_dfts(fdtd_engine* h, uint32_t ts) {
    for (auto b : h->dft) {
        if (ts % b->interval) continue;
        int rc;
        switch (h->prec) {
        case FDTD_PREC_FP32: rc = launch_dft<float>(h, b, ts); break;
        case FDTD_PREC_FP64: rc = launch_dft<double>(h, b, ts); break;
        case FDTD_PREC_FP16: rc = launch_dft<__half>(h, b, ts); break;
        default: rc = launch_dft<__hip_bfloat16>(h, b, ts); break;
        }
        if (rc) return rc;
    }
    HIPCHK(hipGetLastError());
    return FDTD_OK;
}

template <typename S, typename H>
int get_field_impl(fdtd_engine* h, int which, H* out) {
    typedef typename StoreTraits<S>::A A;
    const size_t n = (size_t)3 * h->N[0] * h->N[1] * h->N[2];
    DevBuf tmp;
    int rc = tmp.alloc(n * sizeof(H));
    if (rc) return rc;
    hipLaunchKernelGGL((k_pack<S, H>), dim3(2048), dim3(256), 0, h->stream, (which == FDTD_VOLT ? h->V : h->I).template as<S>(), tmp.as<H>(), h->d, make_scales<A>(h), which);
    HIPCHK(hipGetLastError());
    return stage_d2h(h->stream, out, tmp.p, n * sizeof(H));
}

template <typename S, typename H>
int set_field_impl(fdtd_engine* h, int which, const H* in) {
    typedef typename StoreTraits<S>::A A;
    const size_t n = (size_t)3 * h->N[0] * h->N[1] * h->N[2];
    DevBuf tmp;
    int rc = tmp.alloc(n * sizeof(H));
    if (rc) return rc;
    if ((rc = stage_h2d(h->stream, tmp.p, in, n * sizeof(H)))) return rc;
    hipLaunchKernelGGL((k_unpack<S, H>), dim3(2048), dim3(256), 0, h->stream, (which == FDTD_VOLT ? h->V : h->I).template as<S>(), tmp.as<H>(), h->d, make_scales<A>(h), which);
    HIPCHK(hipStreamSynchronize(h->stream));
    return FDTD_OK;
}

template <typename H>
int field_dispatch_get(fdtd_engine* h, int which, H* out) {
    switch (h->prec) {
    case FDTD_PREC_FP32: return get_field_impl<float, H>(h, which, out);
    case FDTD_PREC_FP64: return get_field_impl<double, H>(h, which, out);
    case FDTD_PREC_FP16: return get_field_impl<__half, H>(h, which, out);
    default: return get_field_impl<__hip_bfloat16, H>(h, which, out);
    }
}
template <typename H>
int field_dispatch_set(fdtd_engine* h, int which, const H* in) {
    switch (h->prec) {
    case FDTD_PREC_FP32: return set_field_impl<float, H>(h, which, in);
    case FDTD_PREC_FP64: return set_field_impl<double, H>(h, which, in);
    case FDTD_PREC_FP16: return set_field_impl<__half, H>(h, which, in);
    default: return set_field_impl<__hip_bfloat16, H>(h, which, in);
    }
}

// element stored at device address -> double, on the host
double decode_elem(const fdtd_engine* h, const void* raw, double scale) {
    switch (h->prec) {
    case FDTD_PREC_FP32: return (double)*(const float*)raw;
    case FDTD_PREC_FP64: return *(const double*)raw;
    case FDTD_PREC_FP16: { __half v; memcpy(&v, raw, 2); return (double)((float)__half2float(v) * (float)scale); }
    default: { __hip_bfloat16 v; memcpy(&v, raw, 2); return (double)((float)__bfloat162float(v) * (float)scale); }
    }
}

} // namespace wipems

using namespace wipems;

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* fdtd_version(void) { return FDTD_VERSION_STR; }

int fdtd_last_error(char* buf, size_t len) {
    if (!buf || !len) return FDTD_ERR_ARG;
    snprintf(buf, len, "%s", g_last_error.c_str());
    return FDTD_OK;
}

int fdtd_device_name(char* buf, size_t len) {
    if (!buf || !len) return fail(FDTD_ERR_ARG, "null buffer");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(FDTD_ERR_NODEVICE, "no HIP device available");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    hipDeviceProp_t p;
    HIPCHK(hipGetDeviceProperties(&p, dev));
    snprintf(buf, len, "%s (%s, %d CUs)", p.name, p.gcnArchName, p.multiProcessorCount);
    return FDTD_OK;
}

int fdtd_create(const fdtd_grid_desc* desc, fdtd_engine** out) {
    if (!desc || !out) return fail(FDTD_ERR_ARG, "null argument");
    *out = nullptr;
    for (int n = 0; n < 3; ++n)
        if (desc->num_lines[n] < 2) return fail(FDTD_ERR_ARG, "num_lines[%d]=%u < 2", n, desc->num_lines[n]);
    if (!(desc->dT > 0)) return fail(FDTD_ERR_ARG, "dT must be > 0");
    if (desc->precision < FDTD_PREC_FP32 || desc->precision > FDTD_PREC_BF16) return fail(FDTD_ERR_ARG, "bad precision %d", desc->precision);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(FDTD_ERR_NODEVICE, "no HIP device available");
    fdtd_engine* h = new fdtd_engine();
    h->prec = desc->precision;
    HIPCHK(hipGetDevice(&h->dev));
    for (int n = 0; n < 3; ++n) h->N[n] = desc->num_lines[n];
    h->dT = desc->dT;
    h->elem = h->prec == FDTD_PREC_FP64 ? 8 : (h->prec == FDTD_PREC_FP32 ? 4 : 2);
    h->acc = h->prec == FDTD_PREC_FP64 ? 8 : 4;
    const long long nzp = ((long long)h->N[2] + 15) / 16 * 16;
    h->d.nx = (int)h->N[0]; h->d.ny = (int)h->N[1]; h->d.nz = (int)h->N[2];
    h->d.sy = nzp;
    h->d.sx = nzp * h->N[1];
    h->d.sc = h->d.sx * h->N[0];
    hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete h; return fail(FDTD_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e)); }
    int rc;
    if ((rc = init_scratch(h->dev))) { delete h; return rc; }
    const size_t fbytes = (size_t)3 * h->d.sc * h->elem;
    if ((rc = h->V.alloc(fbytes)) || (rc = h->I.alloc(fbytes)) || (rc = h->step.alloc(sizeof(unsigned)))) { delete h; return rc; }
    if (hipMemsetAsync(h->V.p, 0, fbytes, h->stream) != hipSuccess || hipMemsetAsync(h->I.p, 0, fbytes, h->stream) != hipSuccess ||
        hipMemsetAsync(h->step.p, 0, sizeof(unsigned), h->stream) != hipSuccess || hipStreamSynchronize(h->stream) != hipSuccess) {
        delete h;
        return fail(FDTD_ERR_HIP, "field initialisation failed");
    }
    *out = h;
    return FDTD_OK;
}

int fdtd_destroy(fdtd_engine* h) {
    if (!h) return fail(FDTD_ERR_ARG, "null handle");
    (void)hipStreamSynchronize(h->stream);
    delete h;
    return FDTD_OK;
}

// per-cell index (uint16 or uint32) in the padded field layout, and the E
// and H tables of 8-wide rows
static int upload_index_table(fdtd_engine* h, const uint32_t* index, bool i16, const std::vector<char>& te,
                              const std::vector<char>& th) {
    const size_t ib = i16 ? 2 : 4;
    int rc;
    if ((rc = h->cf_idx.alloc((size_t)h->d.sc * ib))) return rc;
    if ((rc = h->cf_tabE.alloc(te.size())) || (rc = h->cf_tabH.alloc(th.size()))) return rc;
    HIPCHK(hipMemsetAsync(h->cf_idx.p, 0, h->cf_idx.bytes, h->stream));
    std::vector<char> rows((size_t)h->N[1] * h->N[2] * ib);
    for (uint32_t x = 0; x < h->N[0]; ++x) {
        const uint32_t* src = index + (size_t)x * h->N[1] * h->N[2];
        if (i16) { uint16_t* o = (uint16_t*)rows.data(); for (size_t i = 0; i < (size_t)h->N[1] * h->N[2]; ++i) o[i] = (uint16_t)src[i]; }
        else memcpy(rows.data(), src, rows.size());
        HIPCHK(hipMemcpy2DAsync((char*)h->cf_idx.p + (size_t)x * h->d.sx * ib, (size_t)h->d.sy * ib, rows.data(), (size_t)h->N[2] * ib,
                                (size_t)h->N[2] * ib, h->N[1], hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
    }
    HIPCHK(hipMemcpyAsync(h->cf_tabE.p, te.data(), te.size(), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->cf_tabH.p, th.data(), th.size(), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->cf_vv.alloc(0); h->cf_vi.alloc(0); h->cf_ii.alloc(0); h->cf_iv.alloc(0);
    return FDTD_OK;
}

int fdtd_set_coeffs(fdtd_engine* h, int dtype, const void* vv, const void* vi, const void* ii, const void* iv) {
    if (!h || !vv || !vi || !ii || !iv) return fail(FDTD_ERR_ARG, "null argument");
    if (h->numTS) return fail(FDTD_ERR_STATE, "coefficients must be set before the first timestep");
    int rc;
    if ((rc = upload_dense3(h, dtype, vv, h->cf_vv)) || (rc = upload_dense3(h, dtype, vi, h->cf_vi)) ||
        (rc = upload_dense3(h, dtype, ii, h->cf_ii)) || (rc = upload_dense3(h, dtype, iv, h->cf_iv)))
        return rc;
    h->cf_idx.alloc(0); h->cf_tabE.alloc(0); h->cf_tabH.alloc(0);
    h->cf_gE.alloc(0); h->cf_gH.alloc(0);
    h->cf_mode = CF_FULL;
    invalidate_graph(h);
    return FDTD_OK;
}

int fdtd_set_coeffs_compressed(fdtd_engine* h, const uint32_t* index, int dtype, const void* table, uint32_t n_entries) {
    if (!h || !index || !table || !n_entries) return fail(FDTD_ERR_ARG, "null argument");
    if (h->numTS) return fail(FDTD_ERR_STATE, "coefficients must be set before the first timestep");
    const size_t nd = (size_t)h->N[0] * h->N[1] * h->N[2];
    for (size_t i = 0; i < nd; ++i)
        if (index[i] >= n_entries) return fail(FDTD_ERR_RANGE, "index[%zu]=%u >= n_entries=%u", i, index[i], n_entries);
    std::vector<char> tab;
    int rc = convert_host(h, dtype, table, (size_t)n_entries * 12, tab);
    if (rc) return rc;
    // split into 8-wide E and H rows
    const size_t a = h->acc;
    std::vector<char> te((size_t)n_entries * 8 * a, 0), th((size_t)n_entries * 8 * a, 0);
    for (uint32_t r = 0; r < n_entries; ++r) {
        const char* row = tab.data() + (size_t)r * 12 * a;
        memcpy(te.data() + (size_t)r * 8 * a, row, 3 * a);            // vv
        memcpy(te.data() + ((size_t)r * 8 + 4) * a, row + 3 * a, 3 * a);  // vi
        memcpy(th.data() + (size_t)r * 8 * a, row + 6 * a, 3 * a);    // ii
        memcpy(th.data() + ((size_t)r * 8 + 4) * a, row + 9 * a, 3 * a);  // iv
    }
    const bool i16 = n_entries <= 65536;
    if ((rc = upload_index_table(h, index, i16, te, th))) return rc;
    h->cf_gE.alloc(0); h->cf_gH.alloc(0);
    h->n_entries = n_entries;
    h->cf_mode = i16 ? CF_IDX16 : CF_IDX32;
    invalidate_graph(h);
    return FDTD_OK;
}

int fdtd_set_coeffs_factored(fdtd_engine* h, const uint32_t* index, const float* table, uint32_t n_entries, const float* fv,
                             const float* fc) {
    if (!h || !index || !table || !n_entries || !fv || !fc) return fail(FDTD_ERR_ARG, "null argument");
    if (h->numTS) return fail(FDTD_ERR_STATE, "coefficients must be set before the first timestep");
    if (h->acc != 4) return fail(FDTD_ERR_ARG, "factored coefficients need float arithmetic (FP32, FP16 or BF16 handles)");
    if (n_entries > 65536) return fail(FDTD_ERR_RANGE, "factored tables hold at most 65536 rows (got %u)", n_entries);
    const size_t nd = (size_t)h->N[0] * h->N[1] * h->N[2];
    for (size_t i = 0; i < nd; ++i)
        if (index[i] >= n_entries) return fail(FDTD_ERR_RANGE, "index[%zu]=%u >= n_entries=%u", i, index[i], n_entries);
    const size_t L = (size_t)h->N[0] + h->N[1] + h->N[2];
    for (size_t i = 0; i < 3 * L; ++i)
        if (!std::isfinite(fv[i]) || !std::isfinite(fc[i])) return fail(FDTD_ERR_ARG, "mesh factor %zu is not finite", i);
    std::vector<char> te((size_t)n_entries * 8 * 4, 0), th((size_t)n_entries * 8 * 4, 0);
    for (uint32_t r = 0; r < n_entries; ++r) {
        const float* row = table + (size_t)r * 12;
        memcpy(te.data() + (size_t)r * 32, row, 12);           // vv
        memcpy(te.data() + (size_t)r * 32 + 16, row + 3, 12);  // m_E
        memcpy(th.data() + (size_t)r * 32, row + 6, 12);       // ii
        memcpy(th.data() + (size_t)r * 32 + 16, row + 9, 12);  // m_H
    }
    int rc;
    if ((rc = upload_index_table(h, index, true, te, th))) return rc;
    // factors per component: x lines, y lines (both padded to 16), z lines
    // padded to the field row length; padding holds zeros
    const int gx = (int)((h->N[0] + 15) / 16 * 16), gy = (int)((h->N[1] + 15) / 16 * 16);
    const int gs = gx + gy + (int)h->d.sy;
    std::vector<float> ge((size_t)3 * gs, 0.0f), gh((size_t)3 * gs, 0.0f);
    for (int n = 0; n < 3; ++n) {
        const float* sv = fv + n * L;
        const float* sc = fc + n * L;
        for (uint32_t i = 0; i < h->N[0]; ++i) { ge[(size_t)n * gs + i] = sv[i]; gh[(size_t)n * gs + i] = sc[i]; }
        for (uint32_t i = 0; i < h->N[1]; ++i) { ge[(size_t)n * gs + gx + i] = sv[h->N[0] + i]; gh[(size_t)n * gs + gx + i] = sc[h->N[0] + i]; }
        for (uint32_t i = 0; i < h->N[2]; ++i) {
            ge[(size_t)n * gs + gx + gy + i] = sv[h->N[0] + h->N[1] + i];
            gh[(size_t)n * gs + gx + gy + i] = sc[h->N[0] + h->N[1] + i];
        }
    }
    if ((rc = h->cf_gE.alloc(ge.size() * 4)) || (rc = h->cf_gH.alloc(gh.size() * 4))) return rc;
    HIPCHK(hipMemcpyAsync(h->cf_gE.p, ge.data(), ge.size() * 4, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->cf_gH.p, gh.data(), gh.size() * 4, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->cf_gy = gx; h->cf_gz = gx + gy; h->cf_gs = gs;
    h->n_entries = n_entries;
    h->cf_mode = CF_FACT16;
    invalidate_graph(h);
    return FDTD_OK;
}

int fdtd_set_storage_scale(fdtd_engine* h, const double volt_scale[3], const double curr_scale[3]) {
    if (!h || !volt_scale || !curr_scale) return fail(FDTD_ERR_ARG, "null argument");
    if (h->numTS) return fail(FDTD_ERR_STATE, "scales must be set before the first timestep");
    for (int n = 0; n < 3; ++n) {
        if (!(volt_scale[n] > 0) || !(curr_scale[n] > 0)) return fail(FDTD_ERR_ARG, "scales must be > 0");
        if (h->prec == FDTD_PREC_FP16 || h->prec == FDTD_PREC_BF16) { h->vscale[n] = volt_scale[n]; h->cscale[n] = curr_scale[n]; }
    }
    invalidate_graph(h);
    return FDTD_OK;
}

int fdtd_set_excitation(fdtd_engine* h, const fdtd_excitation* x) {
    if (!h || !x) return fail(FDTD_ERR_ARG, "null argument");
    if (h->numTS) return fail(FDTD_ERR_STATE, "excitation must be set before the first timestep");
    if (!x->length || !x->signal_volt || !x->signal_curr) return fail(FDTD_ERR_ARG, "empty excitation signal");
    const size_t a = h->acc;
    const uint32_t nv = x->volt_count, nc = x->curr_count;
    auto check = [&](uint32_t cnt, const uint32_t* pos, const uint32_t* dir, const double* amp, const uint32_t* del) -> int {
        if (!cnt) return FDTD_OK;
        if (!pos || !dir || !amp || !del) return fail(FDTD_ERR_ARG, "null excitation list");
        for (uint32_t i = 0; i < cnt; ++i) {
            if (dir[i] > 2) return fail(FDTD_ERR_RANGE, "excitation dir %u", dir[i]);
            for (int n = 0; n < 3; ++n)
                if (pos[(size_t)n * cnt + i] >= h->N[n]) return fail(FDTD_ERR_RANGE, "excitation position out of range");
        }
        return FDTD_OK;
    };
    int rc;
    if ((rc = check(nv, x->volt_pos, x->volt_dir, x->volt_amp, x->volt_delay))) return rc;
    if ((rc = check(nc, x->curr_pos, x->curr_dir, x->curr_amp, x->curr_delay))) return rc;
    // one buffer: sigV, sigI, then per list: off(8) comp(4) delay(4) amp(a)
    const size_t L = x->length;
    h->exc_host_off[0].clear();
    h->exc_host_off[1].clear();
    size_t bytes = 2 * L * a + (size_t)(nv + nc) * (8 + 4 + 4 + a) + 64;
    if ((rc = h->exc_buf.alloc(bytes))) return rc;
    std::vector<char> hb(bytes, 0);
    size_t o = 0;
    auto put_vals = [&](const double* src, size_t n) -> size_t {
        size_t at = o;
        for (size_t i = 0; i < n; ++i) {
            if (a == 8) memcpy(hb.data() + o + i * 8, &src[i], 8);
            else { float f = (float)src[i]; memcpy(hb.data() + o + i * 4, &f, 4); }
        }
        o += n * a;
        o = (o + 7) / 8 * 8;
        return at;
    };
    const size_t o_sv = put_vals(x->signal_volt, L);
    const size_t o_si = put_vals(x->signal_curr, L);
    auto put_list = [&](uint32_t cnt, const uint32_t* pos, const uint32_t* dir, const double* amp, const uint32_t* del, ExcDev& ed, size_t o_sig) {
        ed.count = (int)cnt;
        ed.length = (int)L;
        ed.period = (int)x->period;
        ed.sig = (char*)h->exc_buf.p + o_sig;
        if (!cnt) return;
        size_t o_off = o; o += (size_t)cnt * 8;
        size_t o_comp = o; o += (size_t)cnt * 4;
        size_t o_del = o; o += (size_t)cnt * 4;
        o = (o + 7) / 8 * 8;
        std::vector<long long> keys(cnt);
        for (uint32_t i = 0; i < cnt; ++i) {
            long long off = (long long)dir[i] * h->d.sc + cell_off(h, pos[i], pos[cnt + i], pos[2 * (size_t)cnt + i]);
            memcpy(hb.data() + o_off + i * 8, &off, 8);
            int c = (int)dir[i], dl = (int)del[i];
            memcpy(hb.data() + o_comp + i * 4, &c, 4);
            memcpy(hb.data() + o_del + i * 4, &dl, 4);
            keys[i] = off;
        }
        (&ed == &h->exc_v ? h->exc_host_off[0] : h->exc_host_off[1]).assign(keys.begin(), keys.end());
        size_t o_amp = put_vals(amp, cnt);
        std::sort(keys.begin(), keys.end());
        ed.sequential = std::adjacent_find(keys.begin(), keys.end()) != keys.end();
        ed.off = (const long long*)((char*)h->exc_buf.p + o_off);
        ed.comp = (const int*)((char*)h->exc_buf.p + o_comp);
        ed.delay = (const int*)((char*)h->exc_buf.p + o_del);
        ed.amp = (char*)h->exc_buf.p + o_amp;
    };
    // the voltage hook injects the voltage signal, the current hook the current signal
    put_list(nv, x->volt_pos, x->volt_dir, x->volt_amp, x->volt_delay, h->exc_v, o_sv);
    put_list(nc, x->curr_pos, x->curr_dir, x->curr_amp, x->curr_delay, h->exc_c, o_si);
    HIPCHK(hipMemcpyAsync(h->exc_buf.p, hb.data(), bytes, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->has_exc = true;
    h->hooks_dirty = true;
    invalidate_graph(h);
    return FDTD_OK;
}

int fdtd_add_upml(fdtd_engine* h, const uint32_t start[3], const uint32_t num_lines[3], int dtype,
                  const void* vv, const void* vvfo, const void* vvfn, const void* ii, const void* iifo, const void* iifn) {
    if (!h || !start || !num_lines || !vv || !vvfo || !vvfn || !ii || !iifo || !iifn) return fail(FDTD_ERR_ARG, "null argument");
    if (h->numTS) return fail(FDTD_ERR_STATE, "UPML must be added before the first timestep");
    if (h->pml_final) return fail(FDTD_ERR_STATE, "UPML slabs are fixed once stepping has been prepared");
    if (h->pml.count >= MAX_PML_SLABS) return fail(FDTD_ERR_RANGE, "at most %d UPML slabs", MAX_PML_SLABS);
    for (int n = 0; n < 3; ++n)
        if (!num_lines[n] || start[n] + num_lines[n] > h->N[n]) return fail(FDTD_ERR_RANGE, "UPML slab out of range on axis %d", n);
    for (int s = 0; s < h->pml.count; ++s) {
        const PmlSlab& b = h->pml.slab[s];
        bool overlap = true;
        for (int n = 0; n < 3; ++n)
            if (start[n] + num_lines[n] <= (uint32_t)b.s[n] || (uint32_t)(b.s[n] + b.n[n]) <= start[n]) overlap = false;
        if (overlap) return fail(FDTD_ERR_ARG, "UPML slabs overlap");
    }
    PmlSlab b{};
    for (int n = 0; n < 3; ++n) { b.s[n] = (int)start[n]; b.n[n] = (int)num_lines[n]; }
    // z range aligned to one vector of the main arrays (16 bytes of storage:
    // 4 cells FP32, 8 FP16/BF16, 2 FP64), so slab rows are as short as the
    // vectors allow (a z slab of 8 cells: 8 or 12 cells per row, not 16 or 32)
    const int w = (int)(16 / h->elem);
    b.za = b.s[2] / w * w;
    b.nzp = (b.s[2] + b.n[2] - b.za + w - 1) / w * w;
    b.sc = (long long)num_lines[0] * num_lines[1] * b.nzp;
    b.off = h->pml_total;
    b.ioff = h->pml_cells;
    const size_t cnt = (size_t)3 * b.sc;
    const size_t dense = (size_t)3 * num_lines[0] * num_lines[1] * num_lines[2];
    std::vector<char> all((size_t)6 * cnt * h->acc, 0);
    const void* src[6] = {vv, vvfo, vvfn, ii, iifo, iifn};
    std::vector<char> tmp;
    const size_t a = h->acc;
    for (int k = 0; k < 6; ++k) {
        int rc = convert_host(h, dtype, src[k], dense, tmp);
        if (rc) return rc;
        // dense [3][nx][ny][nz] -> padded [3][nx][ny][nzp] at z offset s[2]-za
        const size_t rows = (size_t)3 * num_lines[0] * num_lines[1];
        for (size_t r = 0; r < rows; ++r)
            memcpy(all.data() + ((size_t)k * cnt + r * b.nzp + (b.s[2] - b.za)) * a, tmp.data() + r * num_lines[2] * a, num_lines[2] * a);
    }
    // offsets are in elements per array; each array covers all slabs, comp-major per slab
    h->pml.slab[h->pml.count++] = b;
    h->pml_total += (long long)cnt;
    h->pml_cells += b.sc;
    h->pml_host_coef.push_back(std::move(all));
    h->pml_dirty = true;
    invalidate_graph(h);
    return FDTD_OK;
}

int fdtd_add_ade(fdtd_engine* h, uint32_t count, const uint32_t* pos, int dtype, const void* v_int, const void* v_ext,
                 const void* v_lor, const void* i_int, const void* i_ext, const void* i_lor, uint32_t flags) {
    if (!h || !pos) return fail(FDTD_ERR_ARG, "null argument");
    if (h->numTS) return fail(FDTD_ERR_STATE, "ADE must be added before the first timestep");
    if (!count) return FDTD_OK;
    if ((flags & 1) && (!v_int || !v_ext)) return fail(FDTD_ERR_ARG, "voltage ADE lists missing");
    if ((flags & 2) && (!i_int || !i_ext)) return fail(FDTD_ERR_ARG, "current ADE lists missing");
    if ((flags & 4) && !v_lor) return fail(FDTD_ERR_ARG, "voltage Lorentz list missing");
    if ((flags & 8) && !i_lor) return fail(FDTD_ERR_ARG, "current Lorentz list missing");
    std::vector<long long> off(count);
    for (uint32_t i = 0; i < count; ++i) {
        for (int n = 0; n < 3; ++n)
            if (pos[(size_t)n * count + i] >= h->N[n]) return fail(FDTD_ERR_RANGE, "ADE position out of range");
        off[i] = cell_off(h, pos[i], pos[count + i], pos[2 * (size_t)count + i]);
    }
    AdeHost* a = new AdeHost();
    a->count = count;
    a->flags = flags;
    const size_t n3 = (size_t)3 * count, ab = h->acc;
    int rc;
    if ((rc = a->off.alloc(count * sizeof(long long))) || (rc = a->state.alloc(4 * n3 * ab))) { delete a; return rc; }
    std::vector<char> all(6 * n3 * ab, 0), tmp;
    const void* lists[6] = {v_int, v_ext, v_lor, i_int, i_ext, i_lor};
    for (int k = 0; k < 6; ++k) {
        if (!lists[k]) continue;
        if ((rc = convert_host(h, dtype, lists[k], n3, tmp))) { delete a; return rc; }
        memcpy(all.data() + k * n3 * ab, tmp.data(), n3 * ab);
    }
    // per field, the distinct 9-tuples of the entries' coefficients (the
    // entries of one material share a few); a uint16 row index per entry
    // replaces 36 bytes of list reads per ADE pre-update (FP32)
    bool lists_needed = false;
    for (int w = 0; w < 2; ++w) {
        if (!(flags & (1u << w))) continue;
        std::vector<unsigned short> idx;
        std::vector<char> tab;
        if (h->ade_cmp_mode != 0) {
            std::unordered_map<std::string, unsigned> rows;
            idx.resize(count);
            std::string key(9 * ab, '\0');
            for (uint32_t i = 0; i < count && rows.size() <= 65536; ++i) {
                for (int k = 0; k < 3; ++k)
                    for (int n = 0; n < 3; ++n)
                        memcpy(&key[(k * 3 + n) * ab], all.data() + ((3 * w + k) * n3 + (size_t)n * count + i) * ab, ab);
                auto it = rows.find(key);
                if (it == rows.end()) {
                    if (rows.size() == 65536) { rows.clear(); idx.clear(); break; }
                    it = rows.emplace(key, (unsigned)rows.size()).first;
                    tab.resize(rows.size() * 12 * ab, 0);
                    for (int k = 0; k < 3; ++k) memcpy(&tab[((rows.size() - 1) * 12 + 4 * k) * ab], &key[3 * k * ab], 3 * ab);
                }
                idx[i] = (unsigned short)it->second;
            }
        }
        if (idx.empty()) { lists_needed = true; continue; }
        if ((rc = a->cidx[w].alloc(count * sizeof(unsigned short))) || (rc = a->ctab[w].alloc(tab.size()))) { delete a; return rc; }
        if (hipMemcpyAsync(a->cidx[w].p, idx.data(), count * sizeof(unsigned short), hipMemcpyHostToDevice, h->stream) != hipSuccess ||
            hipMemcpyAsync(a->ctab[w].p, tab.data(), tab.size(), hipMemcpyHostToDevice, h->stream) != hipSuccess ||
            hipStreamSynchronize(h->stream) != hipSuccess) {
            delete a;
            return fail(FDTD_ERR_HIP, "ADE upload failed");
        }
    }
    if (lists_needed && (rc = a->coef.alloc(6 * n3 * ab))) { delete a; return rc; }
    if (hipMemcpyAsync(a->off.p, off.data(), count * sizeof(long long), hipMemcpyHostToDevice, h->stream) != hipSuccess ||
        (lists_needed && hipMemcpyAsync(a->coef.p, all.data(), all.size(), hipMemcpyHostToDevice, h->stream) != hipSuccess) ||
        hipMemsetAsync(a->state.p, 0, a->state.bytes, h->stream) != hipSuccess || hipStreamSynchronize(h->stream) != hipSuccess) {
        delete a;
        return fail(FDTD_ERR_HIP, "ADE upload failed");
    }
    a->host_off = std::move(off);
    h->ade.push_back(a);
    h->hooks_dirty = true;
    invalidate_graph(h);
    return FDTD_OK;
}

int fdtd_set_rlc(fdtd_engine* h, uint32_t count, const uint32_t* pos, const uint32_t* dir, int dtype, const void* coef) {
    if (!h) return fail(FDTD_ERR_ARG, "null handle");
    if (h->numTS) return fail(FDTD_ERR_STATE, "lumped RLC cells must be set before the first timestep");
    delete h->rlc;
    h->rlc = nullptr;
    h->hooks_dirty = true;
    invalidate_graph(h);
    if (!count) return FDTD_OK;
    if (!pos || !dir || !coef) return fail(FDTD_ERR_ARG, "null argument");
    std::vector<long long> off(count);
    std::vector<int> comp(count);
    for (uint32_t i = 0; i < count; ++i) {
        for (int n = 0; n < 3; ++n)
            if (pos[(size_t)n * count + i] >= h->N[n]) return fail(FDTD_ERR_RANGE, "RLC position out of range");
        if (dir[i] > 2) return fail(FDTD_ERR_RANGE, "RLC direction out of range");
        comp[i] = (int)dir[i];
        off[i] = (long long)dir[i] * h->d.sc + cell_off(h, pos[i], pos[count + i], pos[2 * (size_t)count + i]);
    }
    std::vector<long long> sorted(off);
    std::sort(sorted.begin(), sorted.end());
    if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
        return fail(FDTD_ERR_ARG, "two lumped RLC cells on one edge");
    RlcHost* r = new RlcHost();
    const size_t ab = h->acc;
    std::vector<char> c;
    int rc;
    if ((rc = convert_host(h, dtype, coef, (size_t)9 * count, c))) { delete r; return rc; }
    if ((rc = r->off.alloc(count * sizeof(long long))) || (rc = r->comp.alloc(count * sizeof(int))) ||
        (rc = r->coef.alloc(c.size())) || (rc = r->state.alloc((size_t)7 * count * ab))) { delete r; return rc; }
    if (hipMemcpyAsync(r->off.p, off.data(), count * sizeof(long long), hipMemcpyHostToDevice, h->stream) != hipSuccess ||
        hipMemcpyAsync(r->comp.p, comp.data(), count * sizeof(int), hipMemcpyHostToDevice, h->stream) != hipSuccess ||
        hipMemcpyAsync(r->coef.p, c.data(), c.size(), hipMemcpyHostToDevice, h->stream) != hipSuccess ||
        hipMemsetAsync(r->state.p, 0, r->state.bytes, h->stream) != hipSuccess || hipStreamSynchronize(h->stream) != hipSuccess) {
        delete r;
        return fail(FDTD_ERR_HIP, "RLC upload failed");
    }
    r->d.count = (int)count;
    r->d.off = r->off.as<long long>();
    r->d.comp = r->comp.as<int>();
    r->d.c = r->coef.p;
    r->d.il = r->state.p;
    r->d.vd = (char*)r->state.p + (size_t)count * ab;
    r->d.j = (char*)r->state.p + (size_t)4 * count * ab;
    h->rlc = r;
    return FDTD_OK;
}

int fdtd_set_steady_state(fdtd_engine* h, double period, uint32_t count, const uint32_t* pos, const uint32_t* dir) {
    if (!h) return fail(FDTD_ERR_ARG, "null handle");
    if (h->numTS) return fail(FDTD_ERR_STATE, "steady-state detection must be set before the first timestep");
    delete h->ss;
    h->ss = nullptr;
    h->hooks_dirty = true;
    invalidate_graph(h);
    if (!(period > 0)) return FDTD_OK;
    // Operator_Ext_SteadyState::BuildExtension: the period in whole timesteps
    const double pts = round(period / h->dT);
    if (!(pts >= 1) || pts > 1e8) return fail(FDTD_ERR_ARG, "steady-state period must be 1..1e8 timesteps");
    const uint32_t p = (uint32_t)pts;
    if (count > SS_MAX_PROBES) return fail(FDTD_ERR_RANGE, "at most 256 steady-state probes");
    if (count && (!pos || !dir)) return fail(FDTD_ERR_ARG, "null argument");
    std::vector<long long> off(count ? count : 1);
    std::vector<int> comp(count ? count : 1);
    for (uint32_t i = 0; i < count; ++i) {
        for (int n = 0; n < 3; ++n)
            if (pos[(size_t)n * count + i] >= h->N[n]) return fail(FDTD_ERR_RANGE, "steady-state probe position out of range");
        if (dir[i] > 2) return fail(FDTD_ERR_RANGE, "steady-state probe direction out of range");
        comp[i] = (int)dir[i];
        off[i] = (long long)dir[i] * h->d.sc + cell_off(h, pos[i], pos[count + i], pos[2 * (size_t)count + i]);
    }
    int rc;
    if ((rc = ensure_energy_buffers(h))) return rc;
    SsHost* q = new SsHost();
    const size_t nb = count ? count : 1;
    if ((rc = q->off.alloc(nb * sizeof(long long))) || (rc = q->comp.alloc(nb * sizeof(int))) ||
        (rc = q->rec.alloc(nb * 2 * (size_t)p * sizeof(double))) || (rc = q->state.alloc(2 * sizeof(double)))) { delete q; return rc; }
    const double st[2] = {1.0, 0.0};   // last diff starts at 1, last energy at 0
    if (hipMemcpyAsync(q->off.p, off.data(), nb * sizeof(long long), hipMemcpyHostToDevice, h->stream) != hipSuccess ||
        hipMemcpyAsync(q->comp.p, comp.data(), nb * sizeof(int), hipMemcpyHostToDevice, h->stream) != hipSuccess ||
        hipMemsetAsync(q->rec.p, 0, q->rec.bytes, h->stream) != hipSuccess ||
        hipMemcpyAsync(q->state.p, st, sizeof(st), hipMemcpyHostToDevice, h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess) {
        delete q;
        return fail(FDTD_ERR_HIP, "steady-state upload failed");
    }
    q->d.n = (int)count;
    q->d.p = p;
    q->d.off = q->off.as<long long>();
    q->d.comp = q->comp.as<int>();
    q->d.rec = q->rec.as<double>();
    q->d.state = q->state.as<double>();
    h->ss = q;
    return FDTD_OK;
}

int fdtd_get_steady_state_period(fdtd_engine* h, uint32_t* period_ts) {
    if (!h || !period_ts) return fail(FDTD_ERR_ARG, "null argument");
    *period_ts = h->ss ? h->ss->d.p : 0;
    return FDTD_OK;
}

int fdtd_steady_state_last_diff(fdtd_engine* h, double* diff) {
    if (!h || !diff) return fail(FDTD_ERR_ARG, "null argument");
    if (!h->ss) return fail(FDTD_ERR_STATE, "no steady-state detection set");
    double st[2];
    HIPCHK(hipMemcpyAsync(st, h->ss->state.p, sizeof(st), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    *diff = st[0];
    return FDTD_OK;
}

int fdtd_add_mur(fdtd_engine* h, int face, int dtype, const void* k_nyP, const void* k_nyPP, uint32_t start_ts) {
    if (!h || !k_nyP || !k_nyPP) return fail(FDTD_ERR_ARG, "null argument");
    if (h->numTS) return fail(FDTD_ERR_STATE, "Mur faces must be added before the first timestep");
    if (face < 0 || face > 5) return fail(FDTD_ERR_RANGE, "face must be 0..5 (2*axis + top)");
    for (auto m : h->mur)
        if (m->d.ny * 2 + (m->d.line > 0) == face) return fail(FDTD_ERR_ARG, "Mur face %d added twice", face);
    const int ny = face / 2;
    if (h->N[ny] < 2) return fail(FDTD_ERR_RANGE, "Mur face on an axis with fewer than two lines");
    MurHost* m = new MurHost();
    MurDev& d = m->d;
    d.ny = ny; d.nyP = (ny + 1) % 3; d.nyPP = (ny + 2) % 3;
    d.line = (face % 2) ? (int)h->N[ny] - 1 : 0;
    d.shift = (face % 2) ? (int)h->N[ny] - 2 : 1;
    d.na = (int)h->N[d.nyP];
    d.nb = (int)h->N[d.nyPP];
    d.start_ts = start_ts;
    d.sny = ny == 0 ? h->d.sx : (ny == 1 ? h->d.sy : 1);
    const size_t plane = (size_t)d.na * d.nb, ab = h->acc;
    int rc;
    std::vector<char> all(2 * plane * ab), tmp;
    const void* src[2] = {k_nyP, k_nyPP};
    for (int t = 0; t < 2; ++t) {
        if ((rc = convert_host(h, dtype, src[t], plane, tmp))) { delete m; return rc; }
        memcpy(all.data() + t * plane * ab, tmp.data(), plane * ab);
    }
    if ((rc = m->k.alloc(all.size())) || (rc = m->buf.alloc(all.size()))) { delete m; return rc; }
    if (hipMemcpyAsync(m->k.p, all.data(), all.size(), hipMemcpyHostToDevice, h->stream) != hipSuccess ||
        hipMemsetAsync(m->buf.p, 0, m->buf.bytes, h->stream) != hipSuccess || hipStreamSynchronize(h->stream) != hipSuccess) {
        delete m;
        return fail(FDTD_ERR_HIP, "Mur upload failed");
    }
    d.k = m->k.p;
    d.buf = m->buf.p;
    auto it = h->mur.begin();
    while (it != h->mur.end() && (*it)->d.ny * 2 + ((*it)->d.line > 0) < face) ++it;
    h->mur.insert(it, m);
    h->hooks_dirty = true;
    invalidate_graph(h);
    return FDTD_OK;
}

int fdtd_set_tfsf(fdtd_engine* h, int which, uint32_t count, const uint32_t* pos, const uint32_t* comp, const uint32_t* delay,
                  const double* delta, const double* amp, uint32_t max_delay) {
    if (!h) return fail(FDTD_ERR_ARG, "null handle");
    if (h->numTS) return fail(FDTD_ERR_STATE, "TFSF must be set before the first timestep");
    if (!h->has_exc) return fail(FDTD_ERR_STATE, "set the excitation before the TFSF tables");
    if (which != FDTD_VOLT && which != FDTD_CURR) return fail(FDTD_ERR_ARG, "bad field selector");
    TfsfDev& t = which == FDTD_VOLT ? h->tfsf_v : h->tfsf_c;
    DevBuf& buf = which == FDTD_VOLT ? h->tfsf_buf_v : h->tfsf_buf_c;
    t = TfsfDev{};
    h->tfsf_host_off[which == FDTD_VOLT ? 0 : 1].clear();
    h->hooks_dirty = true;
    invalidate_graph(h);
    if (!count) return FDTD_OK;
    if (!pos || !comp || !delay || !delta || !amp) return fail(FDTD_ERR_ARG, "null TFSF list");
    const size_t a = h->acc;
    size_t bytes = (size_t)count * (8 + 4 + 4 + 2 * a) + 64;
    int rc = buf.alloc(bytes);
    if (rc) return rc;
    std::vector<char> hb(bytes, 0);
    size_t o_off = 0, o_comp = (size_t)count * 8, o_del = o_comp + (size_t)count * 4;
    size_t o_dd = (o_del + (size_t)count * 4 + 7) / 8 * 8, o_amp = o_dd + (size_t)count * a;
    // round of each entry = how often its location occurred before it; a
    // stable sort by round keeps list order among repeats of one location
    std::vector<long long> offs(count);
    std::vector<int> round(count);
    std::unordered_map<long long, int> seen;
    int rounds = 0;
    for (uint32_t i = 0; i < count; ++i) {
        for (int n = 0; n < 3; ++n)
            if (pos[(size_t)n * count + i] >= h->N[n]) return fail(FDTD_ERR_RANGE, "TFSF position out of range");
        if (comp[i] > 2) return fail(FDTD_ERR_RANGE, "TFSF component");
        if (delay[i] > max_delay) return fail(FDTD_ERR_RANGE, "TFSF delay beyond max_delay");
        offs[i] = (long long)comp[i] * h->d.sc + cell_off(h, pos[i], pos[count + i], pos[2 * (size_t)count + i]);
        round[i] = seen[offs[i]]++;
        rounds = std::max(rounds, round[i] + 1);
    }
    if (rounds > MAX_TFSF_ROUNDS) return fail(FDTD_ERR_RANGE, "a TFSF location repeats more than %d times", MAX_TFSF_ROUNDS);
    std::vector<uint32_t> order(count);
    for (uint32_t i = 0; i < count; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return round[x] < round[y]; });
    std::vector<long long>& host_off = h->tfsf_host_off[which == FDTD_VOLT ? 0 : 1];
    host_off.resize(count);
    for (uint32_t k = 0; k < count; ++k) {
        const uint32_t i = order[k];
        long long off = offs[i];
        host_off[k] = off;
        int c = (int)comp[i], dl = (int)delay[i];
        memcpy(hb.data() + o_off + k * 8, &off, 8);
        memcpy(hb.data() + o_comp + k * 4, &c, 4);
        memcpy(hb.data() + o_del + k * 4, &dl, 4);
        if (a == 8) { memcpy(hb.data() + o_dd + k * 8, &delta[i], 8); memcpy(hb.data() + o_amp + k * 8, &amp[i], 8); }
        else { float f = (float)delta[i], g = (float)amp[i]; memcpy(hb.data() + o_dd + k * 4, &f, 4); memcpy(hb.data() + o_amp + k * 4, &g, 4); }
    }
    HIPCHK(hipMemcpyAsync(buf.p, hb.data(), bytes, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    char* b = (char*)buf.p;
    t.count = (int)count;
    t.rounds = rounds;
    {
        int k = 0;
        for (int r = 0; r < rounds; ++r) {
            t.rstart[r] = k;
            while (k < (int)count && round[order[k]] == r) ++k;
        }
        t.rstart[rounds] = (int)count;
    }
    t.length = h->exc_v.length;
    t.period = h->exc_v.period;
    t.max_delay = (int)max_delay;
    t.off = (const long long*)(b + o_off);
    t.comp = (const int*)(b + o_comp);
    t.delay = (const int*)(b + o_del);
    t.delta = b + o_dd;
    t.amp = b + o_amp;
    // voltage injection uses the current signal and vice versa (engine_ext_tfsf.cpp:54,145)
    t.sig = which == FDTD_VOLT ? h->exc_c.sig : h->exc_v.sig;
    h->hooks_dirty = true;
    invalidate_graph(h);
    return FDTD_OK;
}

int fdtd_add_probe(fdtd_engine* h, int kind, const uint32_t start[3], const uint32_t stop[3], int norm_dir, uint32_t inside_mask,
                   double weight, uint32_t td_interval, uint32_t fd_interval, const double* freqs, uint32_t n_freq, int* id) {
    if (!h || !start || !stop || !id) return fail(FDTD_ERR_ARG, "null argument");
    if (h->numTS || h->sampled0) return fail(FDTD_ERR_STATE, "probes must be added before the first timestep");
    if (n_freq && (!freqs || !fd_interval)) return fail(FDTD_ERR_ARG, "FD probe needs frequencies and an interval");
    for (int n = 0; n < 3; ++n)
        if (start[n] >= h->N[n] || stop[n] >= h->N[n]) return fail(FDTD_ERR_RANGE, "probe out of range");
    ProbeHost* p = new ProbeHost();
    p->kind = kind;
    p->norm_dir = norm_dir;
    p->weight = weight;
    p->td_interval = td_interval;
    p->fd_interval = fd_interval;
    p->freqs.assign(freqs, freqs + n_freq);
    auto add = [&](int n, uint32_t x, uint32_t y, uint32_t z, int sgn) {
        p->off.push_back((long long)n * h->d.sc + cell_off(h, x, y, z));
        p->comp.push_back(n);
        p->sign.push_back((signed char)sgn);
    };
    if (kind == FDTD_PROBE_VOLTAGE) {
        // Engine_Interface_FDTD::CalcVoltageIntegral, FDTD/engine_interface_fdtd.cpp:200-226
        int ndiff = (start[0] != stop[0]) + (start[1] != stop[1]) + (start[2] != stop[2]);
        if (ndiff != 1) { delete p; return fail(FDTD_ERR_ARG, "voltage probe must be a line along one axis"); }
        for (int n = 0; n < 3; ++n) {
            if (start[n] < stop[n]) {
                uint32_t pos[3] = {start[0], start[1], start[2]};
                for (; pos[n] < stop[n]; ++pos[n]) add(n, pos[0], pos[1], pos[2], +1);
            } else if (start[n] > stop[n]) {
                uint32_t pos[3] = {stop[0], stop[1], stop[2]};
                for (; pos[n] < start[n]; ++pos[n]) add(n, pos[0], pos[1], pos[2], -1);
            }
        }
    } else if (kind == FDTD_PROBE_CURRENT) {
        // ProcessCurrent::CalcIntegral, Common/processcurrent.cpp:93-168
        if (norm_dir < 0 || norm_dir > 2 || start[norm_dir] != stop[norm_dir]) { delete p; return fail(FDTD_ERR_ARG, "current probe needs a valid normal"); }
        const bool si[3] = {(inside_mask & 1) != 0, (inside_mask & 2) != 0, (inside_mask & 4) != 0};
        const bool so[3] = {(inside_mask & 8) != 0, (inside_mask & 16) != 0, (inside_mask & 32) != 0};
        const uint32_t* a = start; const uint32_t* b = stop;
        switch (norm_dir) {
        case 0:
            if (so[0] && si[2]) for (uint32_t i = a[1] + 1; i <= b[1]; ++i) add(1, b[0], i, a[2], +1);
            if (so[0] && so[1]) for (uint32_t i = a[2] + 1; i <= b[2]; ++i) add(2, b[0], b[1], i, +1);
            if (si[0] && so[2]) for (uint32_t i = a[1] + 1; i <= b[1]; ++i) add(1, a[0], i, b[2], -1);
            if (si[0] && si[1]) for (uint32_t i = a[2] + 1; i <= b[2]; ++i) add(2, a[0], a[1], i, -1);
            break;
        case 1:
            if (si[0] && si[1]) for (uint32_t i = a[2] + 1; i <= b[2]; ++i) add(2, a[0], a[1], i, +1);
            if (so[1] && so[2]) for (uint32_t i = a[0] + 1; i <= b[0]; ++i) add(0, i, b[1], b[2], +1);
            if (so[0] && so[1]) for (uint32_t i = a[2] + 1; i <= b[2]; ++i) add(2, b[0], b[1], i, -1);
            if (si[1] && si[2]) for (uint32_t i = a[0] + 1; i <= b[0]; ++i) add(0, i, a[1], a[2], -1);
            break;
        default:
            if (si[1] && si[2]) for (uint32_t i = a[0] + 1; i <= b[0]; ++i) add(0, i, a[1], a[2], +1);
            if (so[0] && si[2]) for (uint32_t i = a[1] + 1; i <= b[1]; ++i) add(1, b[0], i, a[2], +1);
            if (so[1] && so[2]) for (uint32_t i = a[0] + 1; i <= b[0]; ++i) add(0, i, b[1], b[2], -1);
            if (si[0] && so[2]) for (uint32_t i = a[1] + 1; i <= b[1]; ++i) add(1, a[0], i, b[2], -1);
            break;
        }
    } else {
        delete p;
        return fail(FDTD_ERR_ARG, "unknown probe kind %d", kind);
    }
    int rc;
    if (n_freq) {
        if ((rc = p->fd.alloc((size_t)n_freq * 2 * sizeof(double))) || (rc = p->dfreq.alloc((size_t)n_freq * sizeof(double)))) { delete p; return rc; }
        if (hipMemsetAsync(p->fd.p, 0, p->fd.bytes, h->stream) != hipSuccess ||
            hipMemcpyAsync(p->dfreq.p, freqs, (size_t)n_freq * sizeof(double), hipMemcpyHostToDevice, h->stream) != hipSuccess ||
            hipStreamSynchronize(h->stream) != hipSuccess) { delete p; return fail(FDTD_ERR_HIP, "probe upload failed"); }
    }
    *id = (int)h->probes.size();
    h->probes.push_back(p);
    h->probes_dirty = true;
    return FDTD_OK;
}

int fdtd_probe_num_samples(fdtd_engine* h, int id, uint32_t* n) {
    if (!h || !n) return fail(FDTD_ERR_ARG, "null argument");
    if (id < 0 || id >= (int)h->probes.size()) return fail(FDTD_ERR_RANGE, "bad probe id");
    ProbeHost* p = h->probes[id];
    *n = (p->td_interval && h->sampled0) ? h->numTS / p->td_interval + 1 : 0;
    return FDTD_OK;
}

int fdtd_read_probe(fdtd_engine* h, int id, double* t, double* val, uint32_t max_n) {
    if (!h || !t || !val) return fail(FDTD_ERR_ARG, "null argument");
    uint32_t n;
    int rc = fdtd_probe_num_samples(h, id, &n);
    if (rc) return rc;
    n = std::min(n, max_n);
    if (!n) return FDTD_OK;
    std::vector<double> buf((size_t)2 * n);
    HIPCHK(hipMemcpyAsync(buf.data(), h->probes[id]->td.p, buf.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (uint32_t i = 0; i < n; ++i) { t[i] = buf[2 * i]; val[i] = buf[2 * i + 1]; }
    return FDTD_OK;
}

int fdtd_read_probe_fd(fdtd_engine* h, int id, double* re, double* im) {
    if (!h || !re || !im) return fail(FDTD_ERR_ARG, "null argument");
    if (id < 0 || id >= (int)h->probes.size()) return fail(FDTD_ERR_RANGE, "bad probe id");
    ProbeHost* p = h->probes[id];
    const size_t nf = p->freqs.size();
    if (!nf) return FDTD_OK;
    std::vector<double> buf(2 * nf);
    HIPCHK(hipMemcpyAsync(buf.data(), p->fd.p, buf.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (size_t k = 0; k < nf; ++k) { re[k] = buf[2 * k]; im[k] = buf[2 * k + 1]; }
    return FDTD_OK;
}

int fdtd_set_edge_lengths(fdtd_engine* h, const double* primary, const double* dual) {
    if (!h || !primary || !dual) return fail(FDTD_ERR_ARG, "null argument");
    const size_t tot = (size_t)h->N[0] + h->N[1] + h->N[2];
    std::vector<double> buf(2 * tot);
    for (size_t k = 0; k < tot; ++k) {
        if (!(primary[k] >= 0.0) || !(dual[k] >= 0.0)) return fail(FDTD_ERR_ARG, "edge lengths must be finite and >= 0");
        buf[k] = primary[k];
        buf[tot + k] = dual[k];
    }
    int rc;
    if (!h->edge_len.p && (rc = h->edge_len.alloc(buf.size() * sizeof(double)))) return rc;
    HIPCHK(hipMemcpyAsync(h->edge_len.p, buf.data(), buf.size() * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->has_edges = true;
    return FDTD_OK;
}

// dump type and interpolation of a new field box, and the state it needs
static int check_dump(const fdtd_engine* h, int dump_type, int interp, const char* what) {
    if (dump_type < FDTD_DUMP_E || dump_type > FDTD_DUMP_B) return fail(FDTD_ERR_ARG, "dump_type must be FDTD_DUMP_E .. FDTD_DUMP_B");
    if (interp < FDTD_INTERP_NONE || interp > FDTD_INTERP_CELL) return fail(FDTD_ERR_ARG, "interp must be FDTD_INTERP_NONE, _NODE or _CELL");
    if (!h->has_edges) return fail(FDTD_ERR_STATE, "fdtd_set_edge_lengths must be called before %s", what);
    const int need = dump_type == FDTD_DUMP_J ? FDTD_MAT_KAPPA : dump_type == FDTD_DUMP_D ? FDTD_MAT_EPS : dump_type == FDTD_DUMP_B ? FDTD_MAT_MU : -1;
    static const char* const names[3] = {"FDTD_MAT_EPS", "FDTD_MAT_KAPPA", "FDTD_MAT_MU"};
    if (need >= 0 && !h->mat[need].p) return fail(FDTD_ERR_STATE, "this dump type needs fdtd_set_material(%s) before %s", names[need], what);
    return FDTD_OK;
}

int fdtd_set_material(fdtd_engine* h, int kind, const float* values) {
    if (!h || !values) return fail(FDTD_ERR_ARG, "null argument");
    if (kind < FDTD_MAT_EPS || kind > FDTD_MAT_MU) return fail(FDTD_ERR_ARG, "kind must be FDTD_MAT_EPS, FDTD_MAT_KAPPA or FDTD_MAT_MU");
    const size_t nd = (size_t)h->N[0] * h->N[1] * h->N[2];
    DevBuf& buf = h->mat[kind];
    int rc;
    if (!buf.p && (rc = buf.alloc((size_t)3 * h->d.sc * sizeof(float)))) return rc;
    HIPCHK(hipMemsetAsync(buf.p, 0, buf.bytes, h->stream));
    DevBuf dense;
    if ((rc = dense.alloc(nd * sizeof(float)))) return rc;
    for (int n = 0; n < 3; ++n) {
        if ((rc = stage_h2d(h->stream, dense.p, values + (size_t)n * nd, nd * sizeof(float)))) return rc;
        HIPCHK(hipMemcpy2DAsync((char*)buf.p + (size_t)n * h->d.sc * sizeof(float), (size_t)h->d.sy * sizeof(float), dense.p,
                                (size_t)h->N[2] * sizeof(float), (size_t)h->N[2] * sizeof(float), (size_t)h->N[0] * h->N[1],
                                hipMemcpyDeviceToDevice, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
    }
    return FDTD_OK;
}

int fdtd_add_dft_box(fdtd_engine* h, int field, const uint32_t start[3], const uint32_t stop[3],
                     const double* freqs, uint32_t n_freq, uint32_t fd_interval, int* id) {
    if (!h) return fail(FDTD_ERR_ARG, "null handle");
    if (field != FDTD_VOLT && field != FDTD_CURR) return fail(FDTD_ERR_ARG, "field must be FDTD_VOLT (E) or FDTD_CURR (H)");
    return fdtd_add_dft_dump_box(h, field == FDTD_CURR ? FDTD_DUMP_H : FDTD_DUMP_E, FDTD_INTERP_NODE, start, stop, freqs, n_freq, fd_interval, id);
}

int fdtd_add_dft_dump_box(fdtd_engine* h, int dump_type, int interp, const uint32_t start[3], const uint32_t stop[3],
                          const double* freqs, uint32_t n_freq, uint32_t fd_interval, int* id) {
    if (!h || !start || !stop || !freqs || !id) return fail(FDTD_ERR_ARG, "null argument");
    int rc;
    if ((rc = check_dump(h, dump_type, interp, "a DFT box is added"))) return rc;
    if (n_freq < 1 || n_freq > MAX_DFT_FREQ) return fail(FDTD_ERR_RANGE, "n_freq must be 1..%d", MAX_DFT_FREQ);
    if (fd_interval < 1) return fail(FDTD_ERR_ARG, "fd_interval must be >= 1");
    auto b = std::unique_ptr<DftHost>(new DftHost());
    b->type = dump_type;
    b->interp = interp;
    b->nodes = 1;
    for (int n = 0; n < 3; ++n) {
        if (start[n] > stop[n] || stop[n] >= h->N[n]) return fail(FDTD_ERR_RANGE, "DFT box out of the mesh on axis %d", n);
        b->s[n] = start[n];
        b->n[n] = stop[n] - start[n] + 1;
        b->nodes *= b->n[n];
    }
    b->interval = fd_interval;
    b->freqs.assign(freqs, freqs + n_freq);
    const size_t bytes = (size_t)n_freq * 3 * b->nodes * 2 * sizeof(double);
    if ((rc = b->acc.alloc(bytes))) return rc;
    if ((rc = b->dfreq.alloc(n_freq * sizeof(double)))) return rc;
    HIPCHK(hipMemsetAsync(b->acc.p, 0, bytes, h->stream));
    HIPCHK(hipMemcpyAsync(b->dfreq.p, freqs, n_freq * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    *id = (int)h->dft.size();
    h->dft.push_back(b.release());
    return FDTD_OK;
}

int fdtd_add_td_box(fdtd_engine* h, int field, const uint32_t start[3], const uint32_t stop[3], uint32_t interval,
                    fdtd_td_sink sink, void* user, int* id) {
    if (!h) return fail(FDTD_ERR_ARG, "null handle");
    if (field != FDTD_VOLT && field != FDTD_CURR) return fail(FDTD_ERR_ARG, "field must be FDTD_VOLT (E) or FDTD_CURR (H)");
    return fdtd_add_td_dump_box(h, field == FDTD_CURR ? FDTD_DUMP_H : FDTD_DUMP_E, FDTD_INTERP_NODE, start, stop, interval, sink, user, id);
}

int fdtd_add_td_dump_box(fdtd_engine* h, int dump_type, int interp, const uint32_t start[3], const uint32_t stop[3],
                         uint32_t interval, fdtd_td_sink sink, void* user, int* id) {
    if (!h || !start || !stop || !sink || !id) return fail(FDTD_ERR_ARG, "null argument");
    int rc;
    if ((rc = check_dump(h, dump_type, interp, "a dump box is added"))) return rc;
    if (interval < 1) return fail(FDTD_ERR_ARG, "interval must be >= 1");
    auto b = std::unique_ptr<TdHost>(new TdHost());
    b->type = dump_type;
    b->interp = interp;
    b->nodes = 1;
    for (int n = 0; n < 3; ++n) {
        if (start[n] > stop[n] || stop[n] >= h->N[n]) return fail(FDTD_ERR_RANGE, "dump box out of the mesh on axis %d", n);
        b->s[n] = start[n];
        b->n[n] = stop[n] - start[n] + 1;
        b->nodes *= b->n[n];
    }
    if ((b->nodes + 255) / 256 > 0x7fffffffLL) return fail(FDTD_ERR_RANGE, "dump box too large");
    b->interval = interval;
    b->sink = sink;
    b->user = user;
    b->id = (int)h->td.size();
    const size_t bytes = (size_t)3 * b->nodes * sizeof(float);
    if ((rc = b->buf.alloc(bytes))) return rc;
    for (int k = 0; k < 2; ++k) {
        HIPCHK(hipHostMalloc((void**)&b->host[k], bytes, hipHostMallocDefault));
        HIPCHK(hipEventCreateWithFlags(&b->ev[k], hipEventDisableTiming));
    }
    *id = b->id;
    h->td.push_back(b.release());
    return FDTD_OK;
}

int fdtd_td_num_samples(fdtd_engine* h, int id, uint64_t* n) {
    if (!h || !n) return fail(FDTD_ERR_ARG, "null argument");
    if (id < 0 || id >= (int)h->td.size()) return fail(FDTD_ERR_RANGE, "bad dump box id");
    *n = h->td[id]->count;
    return FDTD_OK;
}

int fdtd_dft_num_samples(fdtd_engine* h, int id, uint64_t* n) {
    if (!h || !n) return fail(FDTD_ERR_ARG, "null argument");
    if (id < 0 || id >= (int)h->dft.size()) return fail(FDTD_ERR_RANGE, "bad DFT box id");
    *n = (uint64_t)h->dft[id]->samples;
    return FDTD_OK;
}

int fdtd_read_dft(fdtd_engine* h, int id, double* out) {
    if (!h || !out) return fail(FDTD_ERR_ARG, "null argument");
    if (id < 0 || id >= (int)h->dft.size()) return fail(FDTD_ERR_RANGE, "bad DFT box id");
    DftHost* b = h->dft[id];
    return stage_d2h(h->stream, out, b->acc.p, b->acc.bytes);
}

int fdtd_set_shell_align(fdtd_engine* h, int vectors) {
    if (!h) return fail(FDTD_ERR_ARG, "null handle");
    if (vectors != -1 && vectors != 1 && vectors != 2 && vectors != 4 && vectors != 8)
        return fail(FDTD_ERR_ARG, "vectors must be -1 (auto), 1, 2, 4 or 8");
    h->shell_zalign = vectors;
    invalidate_graph(h);
    return FDTD_OK;
}

int fdtd_set_tile_order(fdtd_engine* h, int band_tiles, int xcd) {
    if (!h) return fail(FDTD_ERR_ARG, "null handle");
    if (band_tiles < -1) return fail(FDTD_ERR_ARG, "band_tiles must be -1 (auto), 0 (whole planes) or a tile count");
    if (xcd < -1) return fail(FDTD_ERR_ARG, "xcd must be -1 (auto), 0 (off) or a chunk width in tiles");
    h->tile_band = band_tiles;
    h->tile_xcd = xcd;
    invalidate_graph(h);
    return FDTD_OK;
}

int fdtd_tile_order(fdtd_engine* h, int* band_tiles, int* xcd) {
    if (!h || !band_tiles || !xcd) return fail(FDTD_ERR_ARG, "null argument");
    const int vz = (int)(16 / h->elem);
    const VecLaunch L = vec_launch(h, vz, false);
    VecRange r = L.r;
    (void)vec_grid(h, L.nq / vz, L.nx, r);
    const long long tiles = (L.nq / vz + 255) / 256;
    *band_tiles = r.bpp && r.tpb < tiles ? r.tpb : 0;
    *xcd = r.xb;
    return FDTD_OK;
}

int fdtd_set_ade_compression(fdtd_engine* h, int mode) {
    if (!h) return fail(FDTD_ERR_ARG, "null handle");
    if (mode < -1 || mode > 1) return fail(FDTD_ERR_ARG, "mode must be -1 (auto), 0 or 1");
    h->ade_cmp_mode = mode;
    return FDTD_OK;
}

int fdtd_ade_compressed(fdtd_engine* h, int order, int* volt, int* curr) {
    if (!h || !volt || !curr) return fail(FDTD_ERR_ARG, "null argument");
    if (order < 0 || order >= (int)h->ade.size()) return fail(FDTD_ERR_RANGE, "bad ADE order");
    *volt = h->ade[order]->ctab[0].p ? 1 : 0;
    *curr = h->ade[order]->ctab[1].p ? 1 : 0;
    return FDTD_OK;
}

int fdtd_set_pml_compression(fdtd_engine* h, int mode) {
    if (!h) return fail(FDTD_ERR_ARG, "null handle");
    if (mode < -1 || mode > 1) return fail(FDTD_ERR_ARG, "mode must be -1 (auto), 0 or 1");
    h->pml_cmp_mode = mode;
    invalidate_graph(h);
    return FDTD_OK;
}

int fdtd_pml_compressed(fdtd_engine* h, int* active) {
    if (!h || !active) return fail(FDTD_ERR_ARG, "null argument");
    int rc = finalize_pml(h);
    if (rc) return rc;
    *active = pml_rows_in_use(h) ? 1 : 0;
    return FDTD_OK;
}

int fdtd_set_pml_split(fdtd_engine* h, int mode) {
    if (!h) return fail(FDTD_ERR_ARG, "null handle");
    if (mode < -1 || mode > 1) return fail(FDTD_ERR_ARG, "mode must be -1 (auto), 0 or 1");
    h->pml_split_mode = mode;
    invalidate_graph(h);
    return FDTD_OK;
}

int fdtd_pml_split(fdtd_engine* h, int* active) {
    if (!h || !active) return fail(FDTD_ERR_ARG, "null argument");
    int rc = finalize_pml(h);
    if (rc) return rc;
    *active = pml_split_in_use(h) ? 1 : 0;
    return FDTD_OK;
}

int fdtd_set_fuse_hooks(fdtd_engine* h, int mode) {
    if (!h) return fail(FDTD_ERR_ARG, "null handle");
    if (mode < -1 || mode > 1) return fail(FDTD_ERR_ARG, "mode must be -1 (auto), 0 or 1");
    h->fuse_hooks_mode = mode;
    h->hooks_dirty = true;
    invalidate_graph(h);
    return FDTD_OK;
}

int fdtd_hooks_fused(fdtd_engine* h, int* active) {
    if (!h || !active) return fail(FDTD_ERR_ARG, "null argument");
    int rc = build_hooks(h);
    if (rc) return rc;
    *active = h->hooks_fused ? 1 : 0;
    return FDTD_OK;
}

int fdtd_set_use_graph(fdtd_engine* h, int enable) {
    if (!h) return fail(FDTD_ERR_ARG, "null handle");
    h->use_graph = enable != 0;
    return FDTD_OK;
}

int fdtd_set_kernel_timing(fdtd_engine* h, int enable) {
    if (!h) return fail(FDTD_ERR_ARG, "null handle");
    h->timing = enable != 0;
    h->t_volt = h->t_curr = 0;
    h->n_volt = h->n_curr = 0;
    return FDTD_OK;
}

int fdtd_get_kernel_times(fdtd_engine* h, double* volt_ms, double* curr_ms, uint32_t* nv, uint32_t* nc) {
    if (!h) return fail(FDTD_ERR_ARG, "null handle");
    if (volt_ms) *volt_ms = h->t_volt;
    if (curr_ms) *curr_ms = h->t_curr;
    if (nv) *nv = h->n_volt;
    if (nc) *nc = h->n_curr;
    return FDTD_OK;
}

static int advance(fdtd_engine* h, uint32_t nsteps);

static int iterate_steps(fdtd_engine* h, uint32_t nsteps);

int fdtd_iterate(fdtd_engine* h, uint32_t nsteps) {
    if (!h) return fail(FDTD_ERR_ARG, "null handle");
    const int rc = iterate_steps(h, nsteps);
    // time-domain samples still in flight reach their sinks before returning
    const int rc2 = drain_tds(h);
    return rc ? rc : rc2;
}

static int iterate_steps(fdtd_engine* h, uint32_t nsteps) {
    int rc = ready_check(h);
    if (rc) return rc;
    if ((rc = finalize_pml(h))) return rc;
    if ((rc = build_hooks(h))) return rc;
    if ((rc = finalize_probes(h, h->numTS + nsteps))) return rc;
    if (!h->sampled0) {
        // ProcessingArray::Process before the loop (openems.cpp:1277-1278): sample at numTS=0
        switch (h->prec) {
        case FDTD_PREC_FP32: rc = launch_step_end<float>(h, 0); break;
        case FDTD_PREC_FP64: rc = launch_step_end<double>(h, 0); break;
        case FDTD_PREC_FP16: rc = launch_step_end<__half>(h, 0); break;
        default: rc = launch_step_end<__hip_bfloat16>(h, 0); break;
        }
        if (rc) return rc;
        if ((rc = launch_dfts(h, 0))) return rc;
        if ((rc = launch_tds(h, 0))) return rc;
        h->sampled0 = true;
    }
    if (!nsteps) return FDTD_OK;
    if (h->dft.empty() && h->td.empty() && !h->ss) return advance(h, nsteps);
    // run to each DFT sample step and each steady-state evaluation step in
    // turn; field boxes accumulate after their step, the steady-state check
    // runs inside its step (one step launched on its own)
    uint32_t left = nsteps;
    while (left) {
        uint64_t next = UINT64_MAX;
        for (auto b : h->dft) next = std::min(next, ((uint64_t)h->numTS / b->interval + 1) * b->interval);
        for (auto b : h->td) next = std::min(next, ((uint64_t)h->numTS / b->interval + 1) * b->interval);
        uint64_t run = std::min<uint64_t>(left, next - h->numTS);
        if (h->ss) {
            const uint64_t p = h->ss->d.p;
            const uint64_t ev = std::max<uint64_t>(2 * p, ((uint64_t)h->numTS + p - 1) / p * p);
            if (ev == h->numTS) {
                h->ss_eval_now = true;
                rc = launch_step(h, false, true, true);
                h->ss_eval_now = false;
                if (rc) return rc;
                HIPCHK(hipGetLastError());
                ++h->numTS;
                --left;
                if ((uint64_t)h->numTS == next && ((rc = launch_dfts(h, h->numTS)) || (rc = launch_tds(h, h->numTS)))) return rc;
                continue;
            }
            run = std::min<uint64_t>(run, ev - h->numTS);
        }
        if ((rc = advance(h, (uint32_t)run))) return rc;
        left -= (uint32_t)run;
        if ((uint64_t)h->numTS == next && ((rc = launch_dfts(h, h->numTS)) || (rc = launch_tds(h, h->numTS)))) return rc;
    }
    return FDTD_OK;
}

// nsteps timesteps with no field-box sampling in between
static int advance(fdtd_engine* h, uint32_t nsteps) {
    int rc;
    if (h->timing) {
        // events for every launch of the two main kernels, harvested after the loop
        const size_t need = (size_t)4 * nsteps;
        while (h->ev.size() < need) {
            hipEvent_t e;
            HIPCHK(hipEventCreate(&e));
            h->ev.push_back(e);
        }
        h->ev_used = 0;
        for (uint32_t s = 0; s < nsteps; ++s)
            if ((rc = launch_step(h, true, s == 0, s + 1 == nsteps))) return rc;
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventSynchronize(h->ev[need - 1]));
        for (uint32_t s = 0; s < nsteps; ++s) {
            float a = 0, b = 0;
            HIPCHK(hipEventElapsedTime(&a, h->ev[4 * s], h->ev[4 * s + 1]));
            HIPCHK(hipEventElapsedTime(&b, h->ev[4 * s + 2], h->ev[4 * s + 3]));
            h->t_volt += a; h->t_curr += b;
        }
        h->n_volt += nsteps; h->n_curr += nsteps;
        h->numTS += nsteps;
        return FDTD_OK;
    }
    if (!h->use_graph) {
        for (uint32_t s = 0; s < nsteps; ++s)
            if ((rc = launch_step(h, false, s == 0, s + 1 == nsteps))) return rc;
        HIPCHK(hipGetLastError());
        h->numTS += nsteps;
        return FDTD_OK;
    }
    const int G = 16;
    uint32_t left = nsteps;
    if (left >= (uint32_t)G) {
        if (!h->gexec) { if ((rc = capture_graph(h, G, &h->graph, &h->gexec))) return rc; }
        while (left >= (uint32_t)G) { HIPCHK(hipGraphLaunch(h->gexec, h->stream)); left -= G; }
    }
    if (left) {
        if (!h->gexec1) { if ((rc = capture_graph(h, 1, &h->graph1, &h->gexec1))) return rc; }
        while (left) { HIPCHK(hipGraphLaunch(h->gexec1, h->stream)); --left; }
    }
    h->numTS += nsteps;
    return FDTD_OK;
}

int fdtd_synchronize(fdtd_engine* h) {
    if (!h) return fail(FDTD_ERR_ARG, "null handle");
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipGetLastError());
    return FDTD_OK;
}

int fdtd_get_num_timesteps(fdtd_engine* h, uint32_t* ts) {
    if (!h || !ts) return fail(FDTD_ERR_ARG, "null argument");
    *ts = h->numTS;
    return FDTD_OK;
}

int fdtd_calc_energy(fdtd_engine* h, double* energy) {
    if (!h || !energy) return fail(FDTD_ERR_ARG, "null argument");
    switch (h->prec) {
    case FDTD_PREC_FP32: return energy_impl<float>(h, energy);
    case FDTD_PREC_FP64: return energy_impl<double>(h, energy);
    case FDTD_PREC_FP16: return energy_impl<__half>(h, energy);
    default: return energy_impl<__hip_bfloat16>(h, energy);
    }
}

int fdtd_field_absmax(fdtd_engine* h, int which, double out[3]) {
    if (!h || !out) return fail(FDTD_ERR_ARG, "null argument");
    if (which != FDTD_VOLT && which != FDTD_CURR) return fail(FDTD_ERR_ARG, "bad field selector");
    switch (h->prec) {
    case FDTD_PREC_FP32: return absmax_impl<float>(h, which, out);
    case FDTD_PREC_FP64: return absmax_impl<double>(h, which, out);
    case FDTD_PREC_FP16: return absmax_impl<__half>(h, which, out);
    default: return absmax_impl<__hip_bfloat16>(h, which, out);
    }
}

int fdtd_calc_energy_noalloc(fdtd_engine* h, double* energy) {
    if (!h || !energy) return fail(FDTD_ERR_ARG, "null argument");
    double* part = scratch_of(h->dev);
    if (!part) return fail(FDTD_ERR_STATE, "no scratch address for device %d", h->dev);
    switch (h->prec) {
    case FDTD_PREC_FP32: return energy_noalloc_impl<float>(h, part, energy);
    case FDTD_PREC_FP64: return energy_noalloc_impl<double>(h, part, energy);
    case FDTD_PREC_FP16: return energy_noalloc_impl<__half>(h, part, energy);
    default: return energy_noalloc_impl<__hip_bfloat16>(h, part, energy);
    }
}

int fdtd_get_field(fdtd_engine* h, int which, int dtype, void* out) {
    if (!h || !out) return fail(FDTD_ERR_ARG, "null argument");
    if (which != FDTD_VOLT && which != FDTD_CURR) return fail(FDTD_ERR_ARG, "bad field selector");
    if (dtype == FDTD_DTYPE_F32) return field_dispatch_get<float>(h, which, (float*)out);
    if (dtype == FDTD_DTYPE_F64) return field_dispatch_get<double>(h, which, (double*)out);
    return fail(FDTD_ERR_ARG, "bad dtype");
}

int fdtd_set_field(fdtd_engine* h, int which, int dtype, const void* in) {
    if (!h || !in) return fail(FDTD_ERR_ARG, "null argument");
    if (which != FDTD_VOLT && which != FDTD_CURR) return fail(FDTD_ERR_ARG, "bad field selector");
    if (which == FDTD_CURR && !h->curr_written) { h->curr_written = true; invalidate_graph(h); }
    if (dtype == FDTD_DTYPE_F32) return field_dispatch_set<float>(h, which, (const float*)in);
    if (dtype == FDTD_DTYPE_F64) return field_dispatch_set<double>(h, which, (const double*)in);
    return fail(FDTD_ERR_ARG, "bad dtype");
}

int fdtd_get_value(fdtd_engine* h, int which, uint32_t n, uint32_t x, uint32_t y, uint32_t z, double* v) {
    if (!h || !v) return fail(FDTD_ERR_ARG, "null argument");
    if (which != FDTD_VOLT && which != FDTD_CURR) return fail(FDTD_ERR_ARG, "bad field selector");
    if (n > 2 || x >= h->N[0] || y >= h->N[1] || z >= h->N[2]) return fail(FDTD_ERR_RANGE, "position out of range");
    const long long off = (long long)n * h->d.sc + cell_off(h, x, y, z);
    char raw[8];
    const DevBuf& F = which == FDTD_VOLT ? h->V : h->I;
    HIPCHK(hipMemcpyAsync(raw, (char*)F.p + off * h->elem, h->elem, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    *v = decode_elem(h, raw, which == FDTD_VOLT ? h->vscale[n] : h->cscale[n]);
    return FDTD_OK;
}

int fdtd_set_value(fdtd_engine* h, int which, uint32_t n, uint32_t x, uint32_t y, uint32_t z, double v) {
    if (!h) return fail(FDTD_ERR_ARG, "null handle");
    if (which != FDTD_VOLT && which != FDTD_CURR) return fail(FDTD_ERR_ARG, "bad field selector");
    if (n > 2 || x >= h->N[0] || y >= h->N[1] || z >= h->N[2]) return fail(FDTD_ERR_RANGE, "position out of range");
    if (which == FDTD_CURR && !h->curr_written) { h->curr_written = true; invalidate_graph(h); }
    const long long off = (long long)n * h->d.sc + cell_off(h, x, y, z);
    char raw[8];
    const double s = which == FDTD_VOLT ? h->vscale[n] : h->cscale[n];
    switch (h->prec) {
    case FDTD_PREC_FP32: { float f = (float)v; memcpy(raw, &f, 4); break; }
    case FDTD_PREC_FP64: memcpy(raw, &v, 8); break;
    case FDTD_PREC_FP16: { __half q = __float2half_rn((float)v * (float)(1.0 / s)); memcpy(raw, &q, 2); break; }
    default: { __hip_bfloat16 q = __float2bfloat16((float)v * (float)(1.0 / s)); memcpy(raw, &q, 2); break; }
    }
    DevBuf& F = which == FDTD_VOLT ? h->V : h->I;
    HIPCHK(hipMemcpyAsync((char*)F.p + off * h->elem, raw, h->elem, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return FDTD_OK;
}

int fdtd_copy_bench(uint64_t bytes, int reps, double* gbps) {
    if (!gbps || reps < 1 || bytes < 4096 || bytes > (1ULL << 40)) return fail(FDTD_ERR_ARG, "bad argument");
    const long long n = (long long)(bytes / 16);
    DevBuf a, b;
    int rc;
    if ((rc = a.alloc((size_t)n * 16)) || (rc = b.alloc((size_t)n * 16))) return rc;
    hipStream_t st;
    HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipMemsetAsync(a.p, 0, a.bytes, st));
    const unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, st, a.as<float4>(), b.as<float4>(), n);   // warm-up
    HIPCHK(hipEventRecord(e0, st));
    for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, st, a.as<float4>(), b.as<float4>(), n);
    HIPCHK(hipEventRecord(e1, st));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(st);
    *gbps = 2.0 * (double)n * 16 * reps / (ms * 1e-3) / 1e9;
    return FDTD_OK;
}

int fdtd_device_bytes(fdtd_engine* h, uint64_t* bytes) {
    if (!h || !bytes) return fail(FDTD_ERR_ARG, "null argument");
    uint64_t b = h->V.bytes + h->I.bytes + h->cf_vv.bytes + h->cf_vi.bytes + h->cf_ii.bytes + h->cf_iv.bytes + h->cf_idx.bytes +
                 h->cf_tabE.bytes + h->cf_tabH.bytes + h->cf_gE.bytes + h->cf_gH.bytes + h->pml_coef.bytes + h->pml_flux.bytes + h->exc_buf.bytes;
    for (auto a : h->ade) b += a->off.bytes + a->coef.bytes + a->state.bytes;
    *bytes = b;
    return FDTD_OK;
}

} // extern "C"
