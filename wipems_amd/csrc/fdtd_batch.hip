// fdtd_batch.hip -- many independent FDTD domains advanced together.
//
// A batch groups engine handles that were configured through the ordinary
// single-domain calls (operator, excitation, UPML, probes).  Each timestep is
// a fixed short sequence of launches for the whole batch instead of one
// sequence per domain:
//
//   k_volt_batch_main   every domain's voltage update over its interior box
//                       (no UPML code), or over the whole domain in the
//                       UPML form where the domain is not split,
//   k_volt_batch_shell  the UPML shells of the split domains,
//   k_excite_batch      the soft sources of every domain,
//   k_curr_batch_main / k_curr_batch_shell, the same for the currents,
//   k_step_end_batch    ++numTS and the probes of every domain (one block each).
//
// Each update grid is the concatenation of the domains' grids; a per-block
// table gives the domain of every block, so uneven grid sizes pack without
// padding and large domains get proportionally more blocks.  The step
// sequence is captured once in a HIP graph and replayed.  Per-cell
// arithmetic is the same device code as the single-domain vector kernels,
// so every domain of a batch gives the same bits as running it alone.
#include "fdtd_internal.h"

using namespace wipems;

namespace {

template <typename A>
struct SimDesc {
    Dims d;
    int cm;                 // CF_* coefficient mode
    int main_pml;           // the main launch runs the UPML form (UPML, not split)
    void* V;
    void* I;
    CoeffPtrs<A> cfe, cfh;
    Scales<A> sc;
    PmlSet ps;
    PmlPtrs<A> pp;
    VecRange rv, rc;            // main-launch ranges (interior box when split)
    long long vbpp, cbpp;       // main-launch blocks per plane
    ShellGrid gv, gc;           // UPML shells (split domains)
    long long vstart, cstart;   // first block of this domain in the main volt / curr grids
    long long vsstart, csstart; // ... and in the shell grids
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

// CM >= 0: every domain of the launch has that coefficient mode; CM < 0:
// per-domain dispatch on D.cm
template <typename S, int CM, bool PML, int VZ>
__device__ __forceinline__ void batch_volt_body(const SimDesc<typename StoreTraits<S>::A>& D, int x, long long q0) {
    S* V = (S*)D.V;
    const S* I = (const S*)D.I;
    if constexpr (CM >= 0) { volt_vec_body<S, CM, PML, VZ>(V, I, D.cfe, D.d, D.sc, D.ps, D.pp, x, q0); return; }
    if (D.cm == CF_FULL) volt_vec_body<S, CF_FULL, PML, VZ>(V, I, D.cfe, D.d, D.sc, D.ps, D.pp, x, q0);
    else if (D.cm == CF_IDX16) volt_vec_body<S, CF_IDX16, PML, VZ>(V, I, D.cfe, D.d, D.sc, D.ps, D.pp, x, q0);
    else volt_vec_body<S, CF_IDX32, PML, VZ>(V, I, D.cfe, D.d, D.sc, D.ps, D.pp, x, q0);
}

template <typename S, int CM, bool PML, int VZ>
__device__ __forceinline__ void batch_curr_body(const SimDesc<typename StoreTraits<S>::A>& D, int x, long long q0) {
    S* I = (S*)D.I;
    const S* V = (const S*)D.V;
    if constexpr (CM >= 0) { curr_vec_body<S, CM, PML, VZ>(I, V, D.cfh, D.d, D.sc, D.ps, D.pp, x, q0); return; }
    if (D.cm == CF_FULL) curr_vec_body<S, CF_FULL, PML, VZ>(I, V, D.cfh, D.d, D.sc, D.ps, D.pp, x, q0);
    else if (D.cm == CF_IDX16) curr_vec_body<S, CF_IDX16, PML, VZ>(I, V, D.cfh, D.d, D.sc, D.ps, D.pp, x, q0);
    else curr_vec_body<S, CF_IDX32, PML, VZ>(I, V, D.cfh, D.d, D.sc, D.ps, D.pp, x, q0);
}

// vector of block b of a domain's main launch: plane r.x0 + local / bpp,
// tile local % bpp over the rows [y0, y1), only vectors within [zlo, zhi)
__device__ __forceinline__ bool batch_main_vector(const VecRange& r, const Dims& d, long long local, long long bpp, int vz,
                                                  int& x, long long& q0) {
    const long long px = local / bpp;
    const long long tile = local - px * bpp;
    q0 = (long long)r.y0 * d.sy + (tile * blockDim.x + threadIdx.x) * vz;
    const int y = (int)(q0 / d.sy);
    const int z0 = (int)(q0 - (long long)y * d.sy);
    x = r.x0 + (int)px;
    return y < r.y1 && z0 >= r.zlo && z0 + vz <= r.zhi;
}

// MAIN_PML: the launch holds only domains whose main launch runs the UPML
// form (the host launches the domains of each kind in a launch of its own)
template <typename S, int CM, bool MAIN_PML>
__global__ void __launch_bounds__(256)
k_volt_batch_main(const SimDesc<typename StoreTraits<S>::A>* __restrict__ sims, const unsigned* __restrict__ blk) {
    typedef typename StoreTraits<S>::A A;
    constexpr int VZ = VecWidth<S>::value;
    const SimDesc<A>& D = sims[blk[blockIdx.x]];
    int x;
    long long q0;
    if (!batch_main_vector(D.rv, D.d, (long long)blockIdx.x - D.vstart, D.vbpp, VZ, x, q0)) return;
    batch_volt_body<S, CM, MAIN_PML, VZ>(D, x, q0);
}

template <typename S, int CM, bool MAIN_PML>
__global__ void __launch_bounds__(256)
k_curr_batch_main(const SimDesc<typename StoreTraits<S>::A>* __restrict__ sims, const unsigned* __restrict__ blk) {
    typedef typename StoreTraits<S>::A A;
    constexpr int VZ = VecWidth<S>::value;
    const SimDesc<A>& D = sims[blk[blockIdx.x]];
    int x;
    long long q0;
    if (!batch_main_vector(D.rc, D.d, (long long)blockIdx.x - D.cstart, D.cbpp, VZ, x, q0)) return;
    batch_curr_body<S, CM, MAIN_PML, VZ>(D, x, q0);
}

// shells of the split domains, vectors of VS cells (shell_vector numbering)
template <typename S, int CM, int VS>
__global__ void __launch_bounds__(256)
k_volt_batch_shell(const SimDesc<typename StoreTraits<S>::A>* __restrict__ sims, const unsigned* __restrict__ blk) {
    typedef typename StoreTraits<S>::A A;
    const SimDesc<A>& D = sims[blk[blockIdx.x]];
    int x;
    long long q0;
    const unsigned t = (unsigned)(((long long)blockIdx.x - D.vsstart) * blockDim.x + threadIdx.x);
    if (!shell_vector(D.gv, t, VS, D.d.sy, x, q0)) return;
    batch_volt_body<S, CM, true, VS>(D, x, q0);
}

template <typename S, int CM, int VS>
__global__ void __launch_bounds__(256)
k_curr_batch_shell(const SimDesc<typename StoreTraits<S>::A>* __restrict__ sims, const unsigned* __restrict__ blk) {
    typedef typename StoreTraits<S>::A A;
    const SimDesc<A>& D = sims[blk[blockIdx.x]];
    int x;
    long long q0;
    const unsigned t = (unsigned)(((long long)blockIdx.x - D.csstart) * blockDim.x + threadIdx.x);
    if (!shell_vector(D.gc, t, VS, D.d.sy, x, q0)) return;
    batch_curr_body<S, CM, true, VS>(D, x, q0);
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

// One update launch of a phase: its grid and per-block domain table.
struct BatchLaunch {
    DevBuf blk;             // uint32 domain index per block
    unsigned blocks = 0;
};

struct fdtd_batch {
    std::vector<fdtd_engine*> eng;
    int prec = FDTD_PREC_FP32;
    hipStream_t stream = nullptr;
    DevBuf sims, vexc, cexc;
    int n_vexc = 0, n_cexc = 0;
    // per phase (0 volt, 1 curr): main launch in the UPML form, main launch
    // without UPML code, shell launch
    BatchLaunch lpml[2], lplain[2], lshell[2];
    int cm = -1;            // coefficient mode shared by every domain, -1 mixed
    int block = 256;
    int split_mode = -1;    // -1 auto (split every domain whose slabs form a shell), 0 never, 1 same as auto
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

// shells of split domains use half-width vectors for 2-byte storage, as the
// single-domain shell launches do
inline int batch_shell_width(size_t elem) { return elem == 2 ? 4 : (int)(16 / elem); }

int upload_table(hipStream_t st, BatchLaunch& L, const std::vector<unsigned>& t) {
    L.blocks = (unsigned)t.size();
    int rc = L.blk.alloc(std::max<size_t>(1, t.size()) * sizeof(unsigned));
    if (rc) return rc;
    if (!t.empty()) HIPCHK(hipMemcpyAsync(L.blk.p, t.data(), t.size() * sizeof(unsigned), hipMemcpyHostToDevice, st));
    return FDTD_OK;
}

template <typename A>
int build_tables(fdtd_batch* b) {
    const int n = (int)b->eng.size();
    std::vector<SimDesc<A>> sd(n);
    std::vector<ExcEntry<A>> ve, ce;
    std::vector<unsigned> tpml[2], tplain[2], tshell[2];
    b->cm = b->eng[0]->cf_mode;
    for (int i = 0; i < n; ++i) {
        fdtd_engine* h = b->eng[i];
        if (h->cf_mode != b->cm) b->cm = -1;
        const int vz = (int)(16 / h->elem);
        const int vs = batch_shell_width(h->elem);
        SimDesc<A>& D = sd[i];
        D = SimDesc<A>{};
        D.d = h->d;
        D.cm = h->cf_mode;
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
        bool split = h->pml.count > 0 && h->shell_ok && b->split_mode != 0;
        if (split) {
            // shell vectors are numbered in 32 bits (at up to twice the count in half-width vectors)
            const VecLaunch L = plan_vec_launch(h->N, h->d.sy, h->pml.count, true, h->ibox, vz, false);
            if (2 * (L.g.nA0 + L.g.nA1 + L.g.nB + L.g.nC) >= (1LL << 31)) split = false;
        }
        if (h->pml.count) {
            A* c = h->pml_coef.as<A>();
            const long long t = h->pml_total;
            D.pp.vv = c; D.pp.vvfo = c + t; D.pp.vvfn = c + 2 * t;
            D.pp.ii = c + 3 * t; D.pp.iifo = c + 4 * t; D.pp.iifn = c + 5 * t;
            D.pp.vflux = h->pml_flux.as<A>(); D.pp.cflux = h->pml_flux.as<A>() + t;
            // coefficient rows as the single-domain engine picks them, with
            // this batch's split in place of the engine's own
            const bool rows = h->pml_cmp && (h->pml_cmp_mode == 1 || (h->pml_cmp_mode < 0 && (h->cf_mode == CF_FULL || split)));
            if (rows) {
                D.pp.vidx = h->pml_vidx.as<unsigned short>(); D.pp.vtab = h->pml_vtab.as<A>();
                D.pp.cidx = h->pml_cidx.as<unsigned short>(); D.pp.ctab = h->pml_ctab.as<A>();
            }
        }
        D.main_pml = h->pml.count > 0 && !split;
        for (int w = 0; w < 2; ++w) {
            const VecLaunch L = plan_vec_launch(h->N, h->d.sy, h->pml.count, split, h->ibox, vz, w == 1, shell_zalign_of(h));
            const ShellGrid g = vs == vz ? L.g : shell_grid_split(L, (int)h->N[2], vz, vz / vs);
            const long long bpp = (L.nq / vz + b->block - 1) / b->block;
            std::vector<unsigned>& tm = D.main_pml ? tpml[w] : tplain[w];
            (w ? D.cstart : D.vstart) = (long long)tm.size();
            (w ? D.rc : D.rv) = L.r;
            (w ? D.cbpp : D.vbpp) = bpp;
            tm.insert(tm.end(), (size_t)(bpp * L.nx), (unsigned)i);
            (w ? D.gc : D.gv) = g;
            (w ? D.csstart : D.vsstart) = (long long)tshell[w].size();
            if (L.shell) {
                const long long nsh = g.nA0 + g.nA1 + g.nB + g.nC;
                tshell[w].insert(tshell[w].end(), (size_t)((nsh + b->block - 1) / b->block), (unsigned)i);
            }
        }
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
    for (int w = 0; w < 2; ++w)
        if (tpml[w].size() >= (1ull << 31) || tplain[w].size() >= (1ull << 31) || tshell[w].size() >= (1ull << 31))
            return fail(FDTD_ERR_ARG, "batch too large: more than 2^31 blocks in one launch");
    int rc;
    for (int w = 0; w < 2; ++w)
        if ((rc = upload_table(b->stream, b->lpml[w], tpml[w])) || (rc = upload_table(b->stream, b->lplain[w], tplain[w])) ||
            (rc = upload_table(b->stream, b->lshell[w], tshell[w])))
            return rc;
    if ((rc = b->sims.alloc(sd.size() * sizeof(SimDesc<A>)))) return rc;
    HIPCHK(hipMemcpyAsync(b->sims.p, sd.data(), sd.size() * sizeof(SimDesc<A>), hipMemcpyHostToDevice, b->stream));
    if ((rc = b->vexc.alloc(std::max<size_t>(1, ve.size()) * sizeof(ExcEntry<A>)))) return rc;
    if ((rc = b->cexc.alloc(std::max<size_t>(1, ce.size()) * sizeof(ExcEntry<A>)))) return rc;
    if (!ve.empty()) HIPCHK(hipMemcpyAsync(b->vexc.p, ve.data(), ve.size() * sizeof(ExcEntry<A>), hipMemcpyHostToDevice, b->stream));
    if (!ce.empty()) HIPCHK(hipMemcpyAsync(b->cexc.p, ce.data(), ce.size() * sizeof(ExcEntry<A>), hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    b->n_vexc = (int)ve.size();
    b->n_cexc = (int)ce.size();
    b->dirty = false;
    b->drop_graph();
    return FDTD_OK;
}

int batch_tables(fdtd_batch* b) {
    return b->prec == FDTD_PREC_FP64 ? build_tables<double>(b) : build_tables<float>(b);
}

// the launches of one update phase (w = 0 volt, 1 curr)
template <typename S, int CM>
void batch_phase(fdtd_batch* b, int w) {
    typedef typename StoreTraits<S>::A A;
    constexpr int VS = sizeof(S) == 2 ? VecWidth<S>::value / 2 : VecWidth<S>::value;
    const SimDesc<A>* sims = b->sims.as<SimDesc<A>>();
    const dim3 blk(b->block);
    if (b->lpml[w].blocks) {
        if (w == 0) hipLaunchKernelGGL((k_volt_batch_main<S, CM, true>), dim3(b->lpml[w].blocks), blk, 0, b->stream, sims, b->lpml[w].blk.as<unsigned>());
        else hipLaunchKernelGGL((k_curr_batch_main<S, CM, true>), dim3(b->lpml[w].blocks), blk, 0, b->stream, sims, b->lpml[w].blk.as<unsigned>());
    }
    if (b->lplain[w].blocks) {
        if (w == 0) hipLaunchKernelGGL((k_volt_batch_main<S, CM, false>), dim3(b->lplain[w].blocks), blk, 0, b->stream, sims, b->lplain[w].blk.as<unsigned>());
        else hipLaunchKernelGGL((k_curr_batch_main<S, CM, false>), dim3(b->lplain[w].blocks), blk, 0, b->stream, sims, b->lplain[w].blk.as<unsigned>());
    }
    if (b->lshell[w].blocks) {
        if (w == 0) hipLaunchKernelGGL((k_volt_batch_shell<S, CM, VS>), dim3(b->lshell[w].blocks), blk, 0, b->stream, sims, b->lshell[w].blk.as<unsigned>());
        else hipLaunchKernelGGL((k_curr_batch_shell<S, CM, VS>), dim3(b->lshell[w].blocks), blk, 0, b->stream, sims, b->lshell[w].blk.as<unsigned>());
    }
}

template <typename S>
void batch_phase_any(fdtd_batch* b, int w) {
    switch (b->cm) {
    case CF_FULL: batch_phase<S, CF_FULL>(b, w); break;
    case CF_IDX16: batch_phase<S, CF_IDX16>(b, w); break;
    case CF_IDX32: batch_phase<S, CF_IDX32>(b, w); break;
    default: batch_phase<S, -1>(b, w); break;
    }
}

template <typename S>
int batch_step(fdtd_batch* b, bool timed, size_t evi) {
    typedef typename StoreTraits<S>::A A;
    const SimDesc<A>* sims = b->sims.as<SimDesc<A>>();
    const int n = (int)b->eng.size();
    if (timed) HIPCHK(hipEventRecord(b->ev[evi], b->stream));
    batch_phase_any<S>(b, 0);
    if (timed) HIPCHK(hipEventRecord(b->ev[evi + 1], b->stream));
    if (b->n_vexc)
        hipLaunchKernelGGL((k_excite_batch<S>), dim3((b->n_vexc + 255) / 256), dim3(256), 0, b->stream, sims, b->vexc.as<ExcEntry<A>>(), b->n_vexc, 0);
    if (timed) HIPCHK(hipEventRecord(b->ev[evi + 2], b->stream));
    batch_phase_any<S>(b, 1);
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

int fdtd_batch_set_split(fdtd_batch* b, int mode) {
    if (!b) return fail(FDTD_ERR_ARG, "null batch");
    if (mode < -1 || mode > 1) return fail(FDTD_ERR_ARG, "mode must be -1 (auto), 0 or 1");
    b->split_mode = mode;
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
