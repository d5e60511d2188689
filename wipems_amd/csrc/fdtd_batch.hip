// fdtd_batch.hip -- many independent FDTD domains advanced together.
//
// A batch groups engine handles that were configured through the ordinary
// single-domain calls (operator, excitation, UPML, probes).  Each timestep is
// three launches for the whole batch instead of three per domain:
//
//   k_volt_batch   every domain's voltage update (UPML fused), then
//   k_excite_batch the soft sources of every domain,
//   k_curr_batch   every domain's current update, then
//   k_step_end_batch  ++numTS and the probes of every domain (one block each).
//
// The update grids are the concatenation of each domain's flat-vector grid
// (x planes x flat chunks); a block finds its domain by binary search over
// the domains' first block index, so uneven grid sizes pack without padding
// and large domains get proportionally more blocks.  The step sequence is
// captured once in a HIP graph and replayed.  Per-cell arithmetic is the same
// device code as the single-domain vector kernels, so every domain of a batch
// gives the same bits as running it alone.
#include "fdtd_internal.h"

using namespace wipems;

namespace {

template <typename A>
struct SimDesc {
    Dims d;
    int cm;                 // CF_* coefficient mode
    int pml;                // 1 if the domain has UPML slabs
    void* V;
    void* I;
    CoeffPtrs<A> cfe, cfh;
    Scales<A> sc;
    PmlSet ps;
    PmlPtrs<A> pp;
    long long vstart, cstart;   // first block of this domain in the volt / curr grids
    long long vbpx, cbpx;       // blocks per x plane
    unsigned* step;
    const ProbeDev* probes;
    int nprobes;
    const long long* e_off;
    const int* e_comp;
    const signed char* e_sign;
    double dT;
};

template <typename A>
struct ExcEntry {
    int sim;
    int comp;
    long long off;          // element offset incl. component
    int delay;
    int length;
    int period;
    A amp;
    const A* sig;
};

template <typename A>
__device__ __forceinline__ int find_sim(const SimDesc<A>* sims, int nsims, long long b, bool curr) {
    int lo = 0, hi = nsims - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const long long st = curr ? sims[mid].cstart : sims[mid].vstart;
        if (st <= b) lo = mid; else hi = mid - 1;
    }
    return lo;
}

template <typename S>
__global__ void __launch_bounds__(256)
k_volt_batch(const SimDesc<typename StoreTraits<S>::A>* __restrict__ sims, int nsims) {
    typedef typename StoreTraits<S>::A A;
    constexpr int VZ = VecWidth<S>::value;
    const long long b = blockIdx.x;
    const SimDesc<A>& D = sims[find_sim(sims, nsims, b, false)];
    const long long local = b - D.vstart;
    const int x = (int)(local / D.vbpx);
    const long long q0 = ((local - (long long)x * D.vbpx) * blockDim.x + threadIdx.x) * VZ;
    S* V = (S*)D.V;
    const S* I = (const S*)D.I;
    if (D.pml) {
        if (D.cm == CF_FULL) volt_vec_body<S, CF_FULL, true, VZ>(V, I, D.cfe, D.d, D.sc, D.ps, D.pp, x, q0);
        else if (D.cm == CF_IDX16) volt_vec_body<S, CF_IDX16, true, VZ>(V, I, D.cfe, D.d, D.sc, D.ps, D.pp, x, q0);
        else volt_vec_body<S, CF_IDX32, true, VZ>(V, I, D.cfe, D.d, D.sc, D.ps, D.pp, x, q0);
    } else {
        if (D.cm == CF_FULL) volt_vec_body<S, CF_FULL, false, VZ>(V, I, D.cfe, D.d, D.sc, D.ps, D.pp, x, q0);
        else if (D.cm == CF_IDX16) volt_vec_body<S, CF_IDX16, false, VZ>(V, I, D.cfe, D.d, D.sc, D.ps, D.pp, x, q0);
        else volt_vec_body<S, CF_IDX32, false, VZ>(V, I, D.cfe, D.d, D.sc, D.ps, D.pp, x, q0);
    }
}

template <typename S>
__global__ void __launch_bounds__(256)
k_curr_batch(const SimDesc<typename StoreTraits<S>::A>* __restrict__ sims, int nsims) {
    typedef typename StoreTraits<S>::A A;
    constexpr int VZ = VecWidth<S>::value;
    const long long b = blockIdx.x;
    const SimDesc<A>& D = sims[find_sim(sims, nsims, b, true)];
    const long long local = b - D.cstart;
    const int x = (int)(local / D.cbpx);
    const long long q0 = ((local - (long long)x * D.cbpx) * blockDim.x + threadIdx.x) * VZ;
    S* I = (S*)D.I;
    const S* V = (const S*)D.V;
    if (D.pml) {
        if (D.cm == CF_FULL) curr_vec_body<S, CF_FULL, true, VZ>(I, V, D.cfh, D.d, D.sc, D.ps, D.pp, x, q0);
        else if (D.cm == CF_IDX16) curr_vec_body<S, CF_IDX16, true, VZ>(I, V, D.cfh, D.d, D.sc, D.ps, D.pp, x, q0);
        else curr_vec_body<S, CF_IDX32, true, VZ>(I, V, D.cfh, D.d, D.sc, D.ps, D.pp, x, q0);
    } else {
        if (D.cm == CF_FULL) curr_vec_body<S, CF_FULL, false, VZ>(I, V, D.cfh, D.d, D.sc, D.ps, D.pp, x, q0);
        else if (D.cm == CF_IDX16) curr_vec_body<S, CF_IDX16, false, VZ>(I, V, D.cfh, D.d, D.sc, D.ps, D.pp, x, q0);
        else curr_vec_body<S, CF_IDX32, false, VZ>(I, V, D.cfh, D.d, D.sc, D.ps, D.pp, x, q0);
    }
}

// Engine_Ext_Excitation::Apply2{Voltages,Current}Impl for every domain
// (engine_ext_excitation.cpp:33-94); edges are distinct within a domain.
template <typename S>
__global__ void k_excite_batch(const SimDesc<typename StoreTraits<S>::A>* __restrict__ sims,
                               const ExcEntry<typename StoreTraits<S>::A>* __restrict__ e, int count, int is_curr) {
    typedef StoreTraits<S> T;
    typedef typename T::A A;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const ExcEntry<A> x = e[i];
    const SimDesc<A>& D = sims[x.sim];
    const int numTS = (int)*D.step;
    int p = numTS + 1;
    if (x.period > 0) p = x.period;
    int k = numTS - x.delay;
    k *= (k > 0);
    k %= p;
    k *= (k < x.length);
    S* fp = (S*)(is_curr ? D.I : D.V) + x.off;
    const A s = is_curr ? D.sc.c[x.comp] : D.sc.v[x.comp];
    A v = T::load(fp, s);
    v = v + x.amp * x.sig[k];
    T::store(fp, v, is_curr ? D.sc.ic[x.comp] : D.sc.iv[x.comp]);
}

template <typename S>
__global__ void k_step_end_batch(const SimDesc<typename StoreTraits<S>::A>* __restrict__ sims, int advance) {
    typedef typename StoreTraits<S>::A A;
    const SimDesc<A>& D = sims[blockIdx.x];
    step_end_body<S>((const S*)D.V, (const S*)D.I, D.step, advance, D.probes, D.nprobes, D.e_off, D.e_comp, D.e_sign, D.sc, D.dT);
}

} // namespace

struct fdtd_batch {
    std::vector<fdtd_engine*> eng;
    int prec = FDTD_PREC_FP32;
    hipStream_t stream = nullptr;
    DevBuf sims, vexc, cexc;
    int n_vexc = 0, n_cexc = 0;
    long long vblocks = 0, cblocks = 0;
    int block = 256;
    bool dirty = true;
    bool use_graph = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph1 = nullptr;
    hipGraphExec_t gexec1 = nullptr;
    bool timing = false;
    std::vector<hipEvent_t> ev;
    double t_volt = 0, t_curr = 0;
    uint32_t launches = 0;
    uint32_t numTS = 0;
    bool sampled0 = false;

    void drop_graph() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        if (gexec1) (void)hipGraphExecDestroy(gexec1);
        if (graph1) (void)hipGraphDestroy(graph1);
        gexec = nullptr; graph = nullptr; gexec1 = nullptr; graph1 = nullptr;
    }
    ~fdtd_batch() {
        drop_graph();
        for (auto e : ev) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

template <typename A>
int build_tables(fdtd_batch* b) {
    const int n = (int)b->eng.size();
    std::vector<SimDesc<A>> sd(n);
    std::vector<ExcEntry<A>> ve, ce;
    long long vb = 0, cb = 0;
    for (int i = 0; i < n; ++i) {
        fdtd_engine* h = b->eng[i];
        const int vz = (int)(16 / h->elem);
        SimDesc<A>& D = sd[i];
        D = SimDesc<A>{};
        D.d = h->d;
        D.cm = h->cf_mode;
        D.pml = h->pml.count > 0;
        D.V = h->V.p;
        D.I = h->I.p;
        if (h->cf_mode == CF_FULL) {
            D.cfe.a = h->cf_vv.as<A>(); D.cfe.b = h->cf_vi.as<A>();
            D.cfh.a = h->cf_ii.as<A>(); D.cfh.b = h->cf_iv.as<A>();
        } else {
            D.cfe.a = h->cf_tabE.as<A>(); D.cfe.idx = h->cf_idx.p;
            D.cfh.a = h->cf_tabH.as<A>(); D.cfh.idx = h->cf_idx.p;
        }
        D.sc = make_scales<A>(h);
        D.ps = h->pml;
        if (h->pml.count) {
            A* c = h->pml_coef.as<A>();
            const long long t = h->pml_total;
            D.pp.vv = c; D.pp.vvfo = c + t; D.pp.vvfn = c + 2 * t;
            D.pp.ii = c + 3 * t; D.pp.iifo = c + 4 * t; D.pp.iifn = c + 5 * t;
            D.pp.vflux = h->pml_flux.as<A>(); D.pp.cflux = h->pml_flux.as<A>() + t;
            if (pml_rows_in_use(h)) {
                D.pp.vidx = h->pml_vidx.as<unsigned short>(); D.pp.vtab = h->pml_vtab.as<A>();
                D.pp.cidx = h->pml_cidx.as<unsigned short>(); D.pp.ctab = h->pml_ctab.as<A>();
            }
        }
        const long long vplane = (long long)h->N[1] * h->d.sy;
        const long long cplane = (long long)(h->N[1] - 1) * h->d.sy;
        D.vbpx = (vplane / vz + b->block - 1) / b->block;
        D.cbpx = (cplane / vz + b->block - 1) / b->block;
        D.vstart = vb;
        D.cstart = cb;
        vb += D.vbpx * h->N[0];
        cb += D.cbpx * (h->N[0] - 1);
        D.step = h->step.as<unsigned>();
        D.probes = h->probe_dev.as<ProbeDev>();
        D.nprobes = (int)h->probes.size();
        D.e_off = h->e_off.as<long long>();
        D.e_comp = h->e_comp.as<int>();
        D.e_sign = h->e_sign.as<signed char>();
        D.dT = h->dT;
        if (h->has_exc) {
            // copy the device excitation lists back once to re-pack them
            for (int w = 0; w < 2; ++w) {
                const ExcDev& ed = w ? h->exc_c : h->exc_v;
                if (!ed.count) continue;
                std::vector<long long> off(ed.count);
                std::vector<int> comp(ed.count), del(ed.count);
                std::vector<A> amp(ed.count);
                HIPCHK(hipMemcpy(off.data(), ed.off, ed.count * sizeof(long long), hipMemcpyDeviceToHost));
                HIPCHK(hipMemcpy(comp.data(), ed.comp, ed.count * sizeof(int), hipMemcpyDeviceToHost));
                HIPCHK(hipMemcpy(del.data(), ed.delay, ed.count * sizeof(int), hipMemcpyDeviceToHost));
                HIPCHK(hipMemcpy(amp.data(), ed.amp, ed.count * sizeof(A), hipMemcpyDeviceToHost));
                for (int k = 0; k < ed.count; ++k) {
                    ExcEntry<A> x{};
                    x.sim = i; x.comp = comp[k]; x.off = off[k]; x.delay = del[k];
                    x.length = ed.length; x.period = ed.period; x.amp = amp[k]; x.sig = (const A*)ed.sig;
                    (w ? ce : ve).push_back(x);
                }
            }
        }
    }
    int rc;
    if ((rc = b->sims.alloc(sd.size() * sizeof(SimDesc<A>)))) return rc;
    HIPCHK(hipMemcpy(b->sims.p, sd.data(), sd.size() * sizeof(SimDesc<A>), hipMemcpyHostToDevice));
    if ((rc = b->vexc.alloc(std::max<size_t>(1, ve.size()) * sizeof(ExcEntry<A>)))) return rc;
    if ((rc = b->cexc.alloc(std::max<size_t>(1, ce.size()) * sizeof(ExcEntry<A>)))) return rc;
    if (!ve.empty()) HIPCHK(hipMemcpy(b->vexc.p, ve.data(), ve.size() * sizeof(ExcEntry<A>), hipMemcpyHostToDevice));
    if (!ce.empty()) HIPCHK(hipMemcpy(b->cexc.p, ce.data(), ce.size() * sizeof(ExcEntry<A>), hipMemcpyHostToDevice));
    b->n_vexc = (int)ve.size();
    b->n_cexc = (int)ce.size();
    b->vblocks = vb;
    b->cblocks = cb;
    b->dirty = false;
    b->drop_graph();
    return FDTD_OK;
}

int batch_tables(fdtd_batch* b) {
    return b->prec == FDTD_PREC_FP64 ? build_tables<double>(b) : build_tables<float>(b);
}

template <typename S>
int batch_step(fdtd_batch* b, bool timed, size_t evi) {
    typedef typename StoreTraits<S>::A A;
    const SimDesc<A>* sims = b->sims.as<SimDesc<A>>();
    const int n = (int)b->eng.size();
    if (timed) HIPCHK(hipEventRecord(b->ev[evi], b->stream));
    hipLaunchKernelGGL((k_volt_batch<S>), dim3((unsigned)b->vblocks), dim3(b->block), 0, b->stream, sims, n);
    if (timed) HIPCHK(hipEventRecord(b->ev[evi + 1], b->stream));
    if (b->n_vexc)
        hipLaunchKernelGGL((k_excite_batch<S>), dim3((b->n_vexc + 255) / 256), dim3(256), 0, b->stream, sims, b->vexc.as<ExcEntry<A>>(), b->n_vexc, 0);
    if (timed) HIPCHK(hipEventRecord(b->ev[evi + 2], b->stream));
    hipLaunchKernelGGL((k_curr_batch<S>), dim3((unsigned)b->cblocks), dim3(b->block), 0, b->stream, sims, n);
    if (timed) HIPCHK(hipEventRecord(b->ev[evi + 3], b->stream));
    if (b->n_cexc)
        hipLaunchKernelGGL((k_excite_batch<S>), dim3((b->n_cexc + 255) / 256), dim3(256), 0, b->stream, sims, b->cexc.as<ExcEntry<A>>(), b->n_cexc, 1);
    hipLaunchKernelGGL((k_step_end_batch<S>), dim3(n), dim3(256), 0, b->stream, sims, 1);
    return FDTD_OK;
}

int batch_step_any(fdtd_batch* b, bool timed, size_t evi) {
    switch (b->prec) {
    case FDTD_PREC_FP32: return batch_step<float>(b, timed, evi);
    case FDTD_PREC_FP64: return batch_step<double>(b, timed, evi);
    case FDTD_PREC_FP16: return batch_step<__half>(b, timed, evi);
    default: return batch_step<__hip_bfloat16>(b, timed, evi);
    }
}

int batch_capture(fdtd_batch* b, int steps, hipGraph_t* g, hipGraphExec_t* ge) {
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    for (int s = 0; s < steps; ++s) {
        int rc = batch_step_any(b, false, 0);
        if (rc) { hipGraph_t dummy; (void)hipStreamEndCapture(b->stream, &dummy); return rc; }
    }
    HIPCHK(hipStreamEndCapture(b->stream, g));
    HIPCHK(hipGraphInstantiate(ge, *g, nullptr, nullptr, 0));
    return FDTD_OK;
}

template <typename S>
void launch_initial_sample(fdtd_batch* b) {
    typedef typename StoreTraits<S>::A A;
    hipLaunchKernelGGL((k_step_end_batch<S>), dim3((unsigned)b->eng.size()), dim3(256), 0, b->stream, b->sims.as<SimDesc<A>>(), 0);
}

} // namespace

extern "C" {

int fdtd_batch_create(fdtd_engine* const* engines, int count, fdtd_batch** out) {
    if (!engines || count < 1 || !out) return fail(FDTD_ERR_ARG, "bad argument");
    *out = nullptr;
    const int prec = engines[0]->prec;
    const uint32_t ts0 = engines[0]->numTS;
    for (int i = 0; i < count; ++i) {
        fdtd_engine* h = engines[i];
        if (!h) return fail(FDTD_ERR_ARG, "null engine %d", i);
        int rc = ready_check(h);
        if (rc) return rc;
        if (h->prec != prec) return fail(FDTD_ERR_ARG, "engines of a batch must share the storage precision");
        if (!h->ade.empty() || h->tfsf_v.count || h->tfsf_c.count)
            return fail(FDTD_ERR_ARG, "engine %d has dispersive or TFSF extensions: not supported in a batch", i);
        if (!h->dft.empty() || !h->td.empty())
            return fail(FDTD_ERR_ARG, "engine %d has field dump boxes: not supported in a batch", i);
        if (!h->mur.empty() || h->rlc || h->ss)
            return fail(FDTD_ERR_ARG, "engine %d has Mur ABC faces, lumped RLC cells or steady-state detection: not supported in a batch", i);
        if (h->exc_v.sequential || h->exc_c.sequential) return fail(FDTD_ERR_ARG, "engine %d has repeated excitation edges", i);
        if (h->cf_mode == CF_FACT16) return fail(FDTD_ERR_ARG, "engine %d has factored coefficients: not supported in a batch", i);
        if (h->numTS != ts0 || h->sampled0 != engines[0]->sampled0) return fail(FDTD_ERR_STATE, "engines of a batch must be at the same timestep");
        for (int j = 0; j < i; ++j)
            if (engines[j] == h) return fail(FDTD_ERR_ARG, "engine %d appears twice", i);
    }
    fdtd_batch* b = new fdtd_batch();
    b->eng.assign(engines, engines + count);
    b->prec = prec;
    b->numTS = ts0;
    b->sampled0 = engines[0]->sampled0;
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete b; return fail(FDTD_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e)); }
    for (auto h : b->eng) {
        int rc;
        if ((rc = finalize_pml(h)) || (rc = fdtd_synchronize(h))) { delete b; return rc; }
    }
    *out = b;
    return FDTD_OK;
}

int fdtd_batch_destroy(fdtd_batch* b) {
    if (!b) return fail(FDTD_ERR_ARG, "null batch");
    (void)hipStreamSynchronize(b->stream);
    delete b;
    return FDTD_OK;
}

int fdtd_batch_set_block(fdtd_batch* b, int threads) {
    if (!b) return fail(FDTD_ERR_ARG, "null batch");
    if (threads != 64 && threads != 128 && threads != 256) return fail(FDTD_ERR_ARG, "block must be 64, 128 or 256 threads");
    b->block = threads;
    b->dirty = true;
    return FDTD_OK;
}

int fdtd_batch_set_use_graph(fdtd_batch* b, int enable) {
    if (!b) return fail(FDTD_ERR_ARG, "null batch");
    b->use_graph = enable != 0;
    return FDTD_OK;
}

int fdtd_batch_set_kernel_timing(fdtd_batch* b, int enable) {
    if (!b) return fail(FDTD_ERR_ARG, "null batch");
    b->timing = enable != 0;
    b->t_volt = b->t_curr = 0;
    b->launches = 0;
    return FDTD_OK;
}

int fdtd_batch_get_kernel_times(fdtd_batch* b, double* volt_ms, double* curr_ms, uint32_t* launches) {
    if (!b) return fail(FDTD_ERR_ARG, "null batch");
    if (volt_ms) *volt_ms = b->t_volt;
    if (curr_ms) *curr_ms = b->t_curr;
    if (launches) *launches = b->launches;
    return FDTD_OK;
}

int fdtd_batch_iterate(fdtd_batch* b, uint32_t nsteps) {
    if (!b) return fail(FDTD_ERR_ARG, "null batch");
    int rc;
    for (auto h : b->eng)
        if (h->numTS != b->numTS) return fail(FDTD_ERR_STATE, "an engine of the batch was stepped on its own");
    // Ordering rule: the probe buffers and the domain table are read by
    // kernels already queued on b->stream, while finalize_probes copies and
    // frees old probe buffers on each member's own stream and batch_tables
    // replaces the table.  Drain b->stream before either can happen.
    bool grow = b->dirty;
    for (auto h : b->eng) grow |= probes_need_realloc(h, b->numTS + nsteps);
    if (grow) HIPCHK(hipStreamSynchronize(b->stream));
    // probe buffers sized for the new horizon (may move them: rebuild tables)
    for (auto h : b->eng) {
        const bool before = h->probes_dirty;
        std::vector<double*> old;
        for (auto p : h->probes) old.push_back(p->td.as<double>());
        if ((rc = finalize_probes(h, b->numTS + nsteps))) return rc;
        bool moved = before;
        for (size_t k = 0; k < h->probes.size(); ++k) moved |= old[k] != h->probes[k]->td.as<double>();
        if (moved) b->dirty = true;
    }
    if (b->dirty && (rc = batch_tables(b))) return rc;
    if (!b->sampled0) {
        switch (b->prec) {
        case FDTD_PREC_FP32: launch_initial_sample<float>(b); break;
        case FDTD_PREC_FP64: launch_initial_sample<double>(b); break;
        case FDTD_PREC_FP16: launch_initial_sample<__half>(b); break;
        default: launch_initial_sample<__hip_bfloat16>(b); break;
        }
        b->sampled0 = true;
        for (auto h : b->eng) h->sampled0 = true;
    }
    if (nsteps) {
        if (b->timing) {
            const size_t need = (size_t)4 * nsteps;
            while (b->ev.size() < need) {
                hipEvent_t e;
                HIPCHK(hipEventCreate(&e));
                b->ev.push_back(e);
            }
            for (uint32_t s = 0; s < nsteps; ++s)
                if ((rc = batch_step_any(b, true, (size_t)4 * s))) return rc;
            HIPCHK(hipGetLastError());
            HIPCHK(hipEventSynchronize(b->ev[need - 1]));
            for (uint32_t s = 0; s < nsteps; ++s) {
                float a = 0, c = 0;
                HIPCHK(hipEventElapsedTime(&a, b->ev[4 * s], b->ev[4 * s + 1]));
                HIPCHK(hipEventElapsedTime(&c, b->ev[4 * s + 2], b->ev[4 * s + 3]));
                b->t_volt += a;
                b->t_curr += c;
            }
            b->launches += nsteps;
        } else if (!b->use_graph) {
            for (uint32_t s = 0; s < nsteps; ++s)
                if ((rc = batch_step_any(b, false, 0))) return rc;
            HIPCHK(hipGetLastError());
        } else {
            const int G = 16;
            uint32_t left = nsteps;
            if (left >= (uint32_t)G) {
                if (!b->gexec && (rc = batch_capture(b, G, &b->graph, &b->gexec))) return rc;
                while (left >= (uint32_t)G) { HIPCHK(hipGraphLaunch(b->gexec, b->stream)); left -= G; }
            }
            if (left) {
                if (!b->gexec1 && (rc = batch_capture(b, 1, &b->graph1, &b->gexec1))) return rc;
                while (left) { HIPCHK(hipGraphLaunch(b->gexec1, b->stream)); --left; }
            }
        }
    }
    b->numTS += nsteps;
    for (auto h : b->eng) h->numTS += nsteps;
    return FDTD_OK;
}

int fdtd_batch_synchronize(fdtd_batch* b) {
    if (!b) return fail(FDTD_ERR_ARG, "null batch");
    HIPCHK(hipStreamSynchronize(b->stream));
    HIPCHK(hipGetLastError());
    return FDTD_OK;
}

} // extern "C"
