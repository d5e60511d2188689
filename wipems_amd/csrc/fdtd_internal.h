// fdtd_internal.h -- host-side state of an engine handle, shared by the
// single-domain (fdtd_engine.hip) and batched (fdtd_batch.hip) C ABI code.
#pragma once
#include "fdtd_device.h"
#include "fdtd_shell.h"

// ---------------------------------------------------------------------------
// error handling
// ---------------------------------------------------------------------------
extern thread_local std::string g_last_error;

inline int fail(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define HIPCHK(expr)                                                                 \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess)                                                        \
            return fail(FDTD_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)


namespace wipems {

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    ~DevBuf() { if (p) (void)hipFree(p); }
    int alloc(size_t n) {
        if (p) { (void)hipFree(p); p = nullptr; bytes = 0; }
        if (n == 0) return FDTD_OK;
        hipError_t e = hipMalloc(&p, n);
        if (e != hipSuccess) { p = nullptr; return fail(FDTD_ERR_NOMEM, "hipMalloc(%zu) failed: %s", n, hipGetErrorString(e)); }
        bytes = n;
        return FDTD_OK;
    }
    template <typename X> X* as() const { return (X*)p; }
};

struct ProbeHost {
    int kind;
    int norm_dir;
    double weight;
    uint32_t td_interval, fd_interval;
    std::vector<double> freqs;
    std::vector<long long> off;
    std::vector<int> comp;
    std::vector<signed char> sign;
    DevBuf td, fd, dfreq;
    long long td_slots = 0;
};

struct DftHost {
    int type = 0;                  // dump type FDTD_DUMP_E .. FDTD_DUMP_B
    int interp = 1;                // FDTD_INTERP_NONE / NODE / CELL
    uint32_t s[3] = {0, 0, 0}, n[3] = {0, 0, 0};
    uint32_t interval = 0;
    long long nodes = 0;
    long long samples = 0;
    std::vector<double> freqs;
    DevBuf acc, dfreq;
};

// time-domain dump box: the snapshot kernel writes one device buffer, which
// is copied to one of two pinned host slots; a slot is handed to the sink
// when it is needed again or when fdtd_iterate returns
struct TdHost {
    int type = 0;                  // dump type FDTD_DUMP_E .. FDTD_DUMP_B
    int interp = 1;                // FDTD_INTERP_NONE / NODE / CELL
    uint32_t s[3] = {0, 0, 0}, n[3] = {0, 0, 0};
    uint32_t interval = 0;
    long long nodes = 0;
    fdtd_td_sink sink = nullptr;
    void* user = nullptr;
    int id = 0;
    DevBuf buf;                    // float [3][nodes]
    float* host[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool pending[2] = {false, false};
    uint32_t pts[2] = {0, 0};
    uint64_t count = 0;
    ~TdHost() {
        for (int k = 0; k < 2; ++k) {
            if (ev[k]) (void)hipEventSynchronize(ev[k]);
            if (host[k]) (void)hipHostFree(host[k]);
            if (ev[k]) (void)hipEventDestroy(ev[k]);
        }
    }
};

struct AdeHost {
    uint32_t count = 0;
    uint32_t flags = 0;
    DevBuf off, coef, state;   // coef: 6 lists [3][count]; state: 4 lists [3][count]
    DevBuf cidx[2], ctab[2];   // coefficient tables of the volt / curr lists (AdeDev::cidx, ctab)
    std::vector<long long> host_off;   // the offsets, kept for the fused hook programs
    DevBuf shared[2];                  // fused hooks: AdeDev::shared masks (volt, curr)
};

// lumped RLC cells: offsets, components, coefficients [9][count], state
struct RlcHost {
    RlcDev d{};
    DevBuf off, comp, coef, state;   // state: il [count], vd [3][count], j [3][count]
};

// steady-state detection: probe list, record ring and the last result
struct SsHost {
    SsDev d{};
    DevBuf off, comp, rec, state;
};

// one Mur ABC face: coefficients and the saved boundary values, [2][na][nb]
struct MurHost {
    MurDev d{};
    DevBuf k, buf;
};

} // namespace wipems

using wipems::DevBuf;
using wipems::ProbeHost;
using wipems::AdeHost;
using wipems::MurHost;
using wipems::RlcHost;
using wipems::SsHost;
using wipems::DftHost;
using wipems::TdHost;

struct fdtd_engine {
    int prec = FDTD_PREC_FP32;
    int dev = 0;
    uint32_t N[3] = {0, 0, 0};
    Dims d{};
    double dT = 0;
    size_t elem = 4;              // bytes per stored field value
    size_t acc = 4;               // bytes per arithmetic value (coefficients)
    hipStream_t stream = nullptr;

    DevBuf V, I;                  // fields (the current ones)
    DevBuf step;                  // unsigned numTS on device
    uint32_t numTS = 0;

    int cf_mode = -1;             // CF_* once set
    DevBuf cf_vv, cf_vi, cf_ii, cf_iv;   // full mode
    DevBuf cf_idx, cf_tabE, cf_tabH;     // compressed and factored modes
    DevBuf cf_gE, cf_gH;                 // factored mode: per-line mesh factors
    int cf_gy = 0, cf_gz = 0, cf_gs = 0; // their y / z offsets and component stride
    uint32_t n_entries = 0;

    double vscale[3] = {1, 1, 1}, cscale[3] = {1, 1, 1};

    // UPML
    PmlSet pml{};
    DevBuf pml_coef;              // 6 arrays over all slabs
    DevBuf pml_flux;              // 2 arrays over all slabs
    std::vector<std::vector<char>> pml_host_coef;   // staged per slab until finalized
    long long pml_total = 0;
    long long pml_cells = 0;      // cells of all slabs (padded layout)
    bool pml_dirty = false;
    bool pml_final = false;
    // compressed UPML coefficients: per cell uint16 rows (volt, curr sides)
    DevBuf pml_vidx, pml_cidx, pml_vtab, pml_ctab;
    bool pml_cmp = false;         // tables built
    int pml_cmp_mode = -1;        // -1 auto (with full main coefficients), 0 off, 1 on
    // slabs forming a shell around the interior box [ibox[n][0], ibox[n][1]):
    // the vector kernels then run the interior without PML code and the
    // shell in a launch of its own
    bool shell_ok = false;
    bool pml_upper = false;       // a slab reaches x = Nx-1, y = Ny-1 or z = Nz-1
    bool curr_written = false;    // the caller wrote current values (SetCurr, fdtd_set_field)
    bool curr_hook_faces = false; // a current-side hook writes a cell of those planes
    int ibox[3][2] = {{0, 0}, {0, 0}, {0, 0}};
    int shell_zalign = -1;        // interior z range aligned to this many vectors: -1 auto (fdtd_set_shell_align)
    int ade_cmp_mode = -1;        // ADE coefficient tables: 0 lists, 1 / -1 tables where they fit (fdtd_set_ade_compression)
    int tile_band = -1;           // tiles per y band of the update launches: -1 auto, 0 whole planes (fdtd_set_tile_order)
    int tile_xcd = -1;            // XCD chunk width in tiles: -1 auto, 0 off
    int pml_split_mode = -1;      // -1 auto (FP16/BF16 storage), 0 off, 1 on

    // excitation
    bool has_exc = false;
    ExcDev exc_v{}, exc_c{};
    DevBuf exc_buf;

    // ADE orders
    std::vector<AdeHost*> ade;

    // Mur ABC faces, kept in ascending face order
    std::vector<MurHost*> mur;

    // lumped RLC cells (one list)
    RlcHost* rlc = nullptr;

    // steady-state detection; ss_eval_now makes the next step evaluate it
    SsHost* ss = nullptr;
    bool ss_eval_now = false;

    // TFSF
    TfsfDev tfsf_v{}, tfsf_c{};
    DevBuf tfsf_buf_v, tfsf_buf_c;

    // fused sparse hooks (k_hooks): host copies of the hook offsets in
    // device order, and the per-location programs of each half step
    std::vector<long long> exc_host_off[2], tfsf_host_off[2];
    HookProg hook[2]{};
    DevBuf hook_buf[2];
    bool hooks_dirty = true;
    bool hooks_fused = false;     // programs built and in use
    int fuse_hooks_mode = -1;     // -1 auto (when no Mur, RLC or steady-state hook interleaves), 0 off, 1 on where possible

    // probes
    std::vector<ProbeHost*> probes;
    DevBuf probe_dev, e_off, e_comp, e_sign;
    bool probes_dirty = true;
    bool sampled0 = false;

    // frequency-domain field boxes
    std::vector<DftHost*> dft;
    // time-domain dump boxes
    std::vector<TdHost*> td;
    DevBuf edge_len;              // [2][Nx+Ny+Nz]: primary, then dual edge lengths (metres)
    bool has_edges = false;
    DevBuf mat[3];                // averaged absolute epsilon, kappa, mu per edge: float [3] padded arrays

    // energy
    DevBuf e_part, e_out;
    int e_blocks = 0;

    // graph
    bool use_graph = true;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    int graph_steps = 0;
    hipGraphExec_t gexec1 = nullptr;
    hipGraph_t graph1 = nullptr;

    // timing
    bool timing = false;
    std::vector<hipEvent_t> ev;
    size_t ev_used = 0;
    double t_volt = 0, t_curr = 0;
    uint32_t n_volt = 0, n_curr = 0;

    ~fdtd_engine() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        if (gexec1) (void)hipGraphExecDestroy(gexec1);
        if (graph1) (void)hipGraphDestroy(graph1);
        for (auto e : ev) (void)hipEventDestroy(e);
        for (auto p : probes) delete p;
        for (auto a : ade) delete a;
        for (auto m : mur) delete m;
        delete rlc;
        delete ss;
        for (auto b : dft) delete b;
        for (auto b : td) delete b;
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace wipems {
void invalidate_graph(fdtd_engine* h);
int finalize_pml(fdtd_engine* h);
void find_shell(fdtd_engine* h);
int compress_pml(fdtd_engine* h);
// split interior / shell launches of the vector kernels (see fdtd_set_pml_split):
// by default where the interior holds at least half of the cells, for
// 2-byte and 4-byte storage; FP64 keeps one launch.  Measured per step on
// MI355X with the interior's z range aligned (shell_zalign_of): 1537x1537x769
// dipole FP32 37.8 (one launch) -> 28.3 ms, dispersive 385x385x193 FP32
// 0.957 -> 0.828 ms (FP16 0.793 -> 0.669), graded 257^3 FP32 with full
// coefficients 0.490 -> 0.473 ms (FP16 0.465 -> 0.326).
inline bool pml_split_in_use(const fdtd_engine* h) {
    if (!h->shell_ok || h->pml_split_mode == 0) return false;
    if (h->pml_split_mode == 1) return true;
    if (h->elem == 8) return false;
    long long inner = 1;
    for (int n = 0; n < 3; ++n) inner *= h->ibox[n][1] - h->ibox[n][0];
    return (double)inner >= 0.5 * (double)h->N[0] * h->N[1] * h->N[2];
}

// UPML coefficient rows: by default with full main coefficients, and with
// the shell split (its scattered vectors gain more from 2-byte rows than the
// extra gathers cost: 1537x1537x769 dipole 16.16 -> 15.27 ms per voltage
// update); an index-form single launch keeps the full slab arrays
inline bool pml_rows_in_use(const fdtd_engine* h) {
    return h->pml_cmp && (h->pml_cmp_mode == 1 || (h->pml_cmp_mode < 0 && (h->cf_mode == CF_FULL || pml_split_in_use(h))));
}
// z alignment (in vectors) of the interior launch of the shell split.
// Auto, measured on MI355X: 4 vectors (64 bytes) for 4-byte storage
// (1537x1537x769 dipole 15.07 -> 14.41 ms per voltage update, 14.63 ->
// 13.97 ms per current update; dispersive 385x385x193 0.364 -> 0.340 and
// 0.354 -> 0.319 ms; 8 vectors is slower than 4), none for 2-byte storage,
// whose half-width shell vectors make every cell moved into the shell launch
// dearer (dispersive FP16 0.259 -> 0.298 ms at 4 vectors, graded 257^3 FP16
// 0.157 -> 0.179 ms).
inline int shell_zalign_of(const fdtd_engine* h) {
    if (h->shell_zalign > 0) return h->shell_zalign;
    return h->elem == 4 ? 4 : 1;
}

// Tile order of the main update launches (see VecRange): tiles per y band
// (0: whole planes) and the XCD chunk width in tiles (0: off).  Auto, from
// measurements on MI355X: bands of 128 tiles with one contiguous sub-band
// per XCD (chunks of 16 tiles) for 4-byte storage without UPML on planes of
// at least 256 tiles (512^3: 0.863 -> 0.844 ms per voltage, 0.907 -> 0.881
// ms per current update, fetched bytes 1.10 -> 1.006 of the algorithmic);
// whole planes in block order otherwise (1537x1537x769 dipole: every band
// size from 584 down to 97 tiles, with and without XCD sub-bands, 1-12%
// slower; 512^3 with 2-byte storage 2-3% slower).
inline bool tile_auto_banded(const fdtd_engine* h, unsigned tiles) {
    return h->elem == 4 && h->pml.count == 0 && tiles >= 256;
}
inline int tile_band_of(const fdtd_engine* h, unsigned tiles) {
    if (h->tile_band > 0) return (int)std::min<unsigned>((unsigned)h->tile_band, tiles);
    if (h->tile_band < 0 && tile_auto_banded(h, tiles)) return 128;
    return 0;
}
inline int tile_xcd_of(const fdtd_engine* h, unsigned tiles) {
    if (h->tile_xcd >= 0) return h->tile_xcd;
    return h->tile_band < 0 && tile_auto_banded(h, tiles) ? 16 : 0;
}
bool probes_need_realloc(const fdtd_engine* h, uint32_t upto_ts);
int finalize_probes(fdtd_engine* h, uint32_t upto_ts);
int ready_check(fdtd_engine* h);

template <typename A>
Scales<A> make_scales(const fdtd_engine* h) {
    Scales<A> s;
    for (int n = 0; n < 3; ++n) {
        s.v[n] = (A)h->vscale[n]; s.iv[n] = (A)(1.0 / h->vscale[n]);
        s.c[n] = (A)h->cscale[n]; s.ic[n] = (A)(1.0 / h->cscale[n]);
    }
    return s;
}
} // namespace wipems
