// operator.cpp -- host-side equivalent-circuit operator builder (C ABI in
// include/wipems_operator.h).
//
// It restates openEMS's Cartesian Operator build for geometry made of boxes
// and spheres (the CSXCAD stand-in), so that the device engine can be fed the
// same vv/vi/ii/iv arrays and extension tables the reference engine would use.
// The template parameter F is the reference's FDTD_FLOAT: float reproduces the
// reference's rounding points (EC_* and the coefficient arrays are float),
// double gives the FP64 oracle operator.
//
// Compiled with -ffp-contract=off so that double expressions round like the
// reference's (baseline x86-64, no FMA).
#include <algorithm>
#include <array>
#include <atomic>
#include <condition_variable>
#include <functional>
#include <memory>
#include <mutex>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <type_traits>
#include <exception>
#include <limits>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include <unistd.h>

#include "wipems_operator.h"

namespace {

const double EPS0 = 8.85418781762e-12;  // tools/constants.h:22
const double MUE0 = 1.256637062e-6;     // tools/constants.h:23
const double C0 = 299792458;            // tools/constants.h:24
const double Z0 = 376.730313461;        // tools/constants.h:25
const double PI_ = 3.141592653589793238462643383279;  // tools/constants.h:26

thread_local std::string g_err;
int fail(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

enum PropKind { P_MATERIAL = 1, P_METAL = 2, P_LUMPED = 4, P_EXCITATION = 8 };

struct Property {
    int kind = 0;
    bool lorentz = false;
    double eps = 1, kappa = 0, mue = 1, sigma = 0;
    double density = 0;       // kg/m^3, for SAR (CSPropMaterial's Density)
    double eps_plasma_f = 0, eps_relax_t = 0, eps_lor_f = 0;
    int ny = 0;
    double R = -1, C = -1;
    double L = -1;            // inductance (lumped RLC); < 0: unset
    int le_type = 0;          // 0 parallel, 1 series (CSPropLumpedElement::LEtype)
    bool caps = false;
    int exc_type = 0;
    double exc_val[3] = {0, 0, 0};
    double delay = 0;
};

// Worker threads kept for the length of one build (PoolScope): a build
// calls parallel_x several times per x plane, and starting threads per call
// cost a third of the streamed build.  run(n, fn) calls fn(0..n-1) on the
// workers and the calling thread and returns when all calls are done; the
// first exception of any call is rethrown there once every worker has
// stopped.  A call made from inside fn runs serially on that thread.
class WorkerPool {
public:
    explicit WorkerPool(int workers) {
        for (int t = 0; t < workers; ++t) th_.emplace_back([this] { loop(); });
    }
    ~WorkerPool() {
        {
            std::lock_guard<std::mutex> l(m_);
            stop_ = true;
        }
        work_.notify_all();
        for (auto& t : th_) t.join();
    }
    size_t workers() const { return th_.size(); }
    static bool inside() { return in_pool(); }
    void run(uint32_t n, const std::function<void(uint32_t)>& fn) {
        {
            std::lock_guard<std::mutex> l(m_);
            job_ = &fn;
            count_ = n;
            next_.store(0);
            busy_ = (int)th_.size();
            ++gen_;
        }
        work_.notify_all();
        drain();
        std::unique_lock<std::mutex> l(m_);
        done_.wait(l, [&] { return busy_ == 0; });
        job_ = nullptr;
        if (err_) {
            std::exception_ptr e = err_;
            err_ = nullptr;
            std::rethrow_exception(e);
        }
    }

private:
    static bool& in_pool() {
        static thread_local bool v = false;
        return v;
    }
    void drain() {
        in_pool() = true;
        try {
            for (uint32_t x; (x = next_.fetch_add(1)) < count_;) (*job_)(x);
        } catch (...) {
            std::lock_guard<std::mutex> l(m_);
            if (!err_) err_ = std::current_exception();
            next_.store(count_);
        }
        in_pool() = false;
    }
    void loop() {
        uint64_t seen = 0;
        for (;;) {
            std::unique_lock<std::mutex> l(m_);
            work_.wait(l, [&] { return stop_ || gen_ != seen; });
            if (stop_) return;
            seen = gen_;
            l.unlock();
            drain();
            l.lock();
            if (--busy_ == 0) done_.notify_one();
        }
    }
    std::vector<std::thread> th_;
    std::mutex m_;
    std::condition_variable work_, done_;
    const std::function<void(uint32_t)>* job_ = nullptr;
    uint32_t count_ = 0;
    std::atomic<uint32_t> next_{0};
    int busy_ = 0;
    uint64_t gen_ = 0;
    bool stop_ = false;
    std::exception_ptr err_;
};

struct Primitive {
    int shape = 0;            // 0 box, 1 sphere
    double a[3] = {0, 0, 0};  // box min / sphere centre
    double b[3] = {0, 0, 0};  // box max / (radius, -, -)
    double bb0[3], bb1[3];    // bounding box
    int priority = 0;
    int prop = 0;
    // CSPrimBox::IsInside is inclusive on both faces; spheres use dist <= r
    bool inside(const double c[3]) const {
        if (shape == 0) {
            for (int n = 0; n < 3; ++n)
                if (c[n] < a[n] || c[n] > b[n]) return false;
            return true;
        }
        double d2 = 0;
        for (int n = 0; n < 3; ++n) { double t = c[n] - a[n]; d2 += t * t; }
        return d2 <= b[0] * b[0];
    }
};

template <typename F>
struct Store {
    std::vector<F> EC[4][3];   // C, G, L, R per direction, as stored by the reference (FDTD_FLOAT)
    std::vector<F> co[4][3];   // vv, vi, ii, iv per direction
    void clear() {
        for (int k = 0; k < 4; ++k)
            for (int n = 0; n < 3; ++n) { std::vector<F>().swap(EC[k][n]); std::vector<F>().swap(co[k][n]); }
    }
};

}  // namespace

struct wop_operator {
    std::vector<double> lines[3];
    uint32_t N[3] = {0, 0, 0};
    double gridDelta = 1;
    int BC[6] = {0, 0, 0, 0, 0, 0};
    uint32_t pml_size[6] = {0, 0, 0, 0, 0, 0};
    double ts_factor = 1;
    double forced_dT = 0;
    int nthreads = 0;

    std::vector<Property> props;
    std::vector<Primitive> prims;   // kept sorted by descending priority (stable)

    int exc_sig_type = WOP_EXC_GAUSS;
    double f0 = 0, fc = 0;

    uint32_t ring = 0;              // plane slots of the streamed build (0: full arrays)
    std::vector<uint32_t> ring_slot;    // x % ring
    bool streamed = false;          // last build kept no full coefficient arrays
    bool store_mat[3] = {false, false, false};  // epsilon, kappa, mu kept for dumps
    std::vector<float> mat[3][3];   // [type][n] averaged material per edge, x-major cell order
    std::vector<double> stream_table;   // its unique coefficient rows [n][12]

    // ---- results ----
    bool built = false;
    int verbose = 0;
    int prec = 0;
    double dT = 0;
    Store<float> sf;                       // float build results
    Store<double> sd;                      // double build results
    template <typename F> Store<F>& store();
    double coef_at(int k, int n, size_t i) const { return prec == 0 ? (double)sf.co[k][n][i] : sd.co[k][n][i]; }
    std::vector<double> sig_v, sig_i;
    uint32_t exc_length = 0, nyquist = 0, max_step = 0;

    struct Exc { uint32_t pos[3]; uint32_t dir; double amp; uint32_t delay; };
    std::vector<Exc> exc_v, exc_i;

    // plane-wave (TFSF) source: a box and the wave's direction, E vector and
    // frequency for the numerical phase velocity (0: vacuum speed of light)
    struct PlaneWave { bool set = false; double a[3], b[3], k[3], E[3], f = 0; } pw;
    struct TfsfEntry { uint32_t pos[3]; uint32_t comp; uint32_t delay; double delta, amp; };
    std::vector<TfsfEntry> tfsf_v, tfsf_i;
    uint32_t tfsf_max_delay = 0;

    struct Upml {
        uint32_t start[3], n[3];
        // vv, vvfo, vvfn, ii, iifo, iifn [3][n], in the build's precision
        // (a float build keeps af, a double build a)
        std::vector<double> a[6];
        std::vector<float> af[6];
    };
    std::vector<Upml> upml;

    // Mur ABC faces (face = 2*ny + top): first-order coefficients per
    // (nyP, nyPP) position of the boundary plane, and the step from which
    // the boundary is active
    double mur_vph[6] = {0, 0, 0, 0, 0, 0};
    struct Mur {
        int face = 0;
        uint32_t na = 0, nb = 0;
        std::vector<double> kP, kPP;       // [na][nb]
        uint32_t start_ts = 0;
    };
    std::vector<Mur> mur;

    // lumped RLC cells (Operator_Ext_LumpedRLC): position, direction and the
    // coefficient lists ilv, i2v, vv2, vj1, vj2, vvd, ib0, b1, b2
    struct Rlc {
        std::vector<uint32_t> pos[3];
        std::vector<uint32_t> dir;
        std::vector<double> c[9];
    } rlc;

    struct Ade {
        std::vector<uint32_t> pos[3];
        std::vector<double> lists[6];      // v_int, v_ext, v_lor, i_int, i_ext, i_lor [3][count]
        uint32_t flags = 0;
    };
    std::vector<Ade> ade;

    // ---- mesh helpers (FDTD/operator.cpp:136-242) ----
    // storage index of a cell; with ring > 0 (streamed build) x planes share
    // `ring` plane slots
    size_t idx(uint32_t x, uint32_t y, uint32_t z) const { return ((size_t)(ring ? ring_slot[x] : x) * N[1] + y) * N[2] + z; }
    size_t nd() const { return (size_t)N[0] * N[1] * N[2]; }
    // material arrays of the flagged types (Operator::InitDataStorage,
    // FDTD/operator.cpp:850-880); the others are released
    void InitMaterialStorage() {
        for (int k = 0; k < 3; ++k)
            for (int n = 0; n < 3; ++n) {
                if (store_mat[k]) mat[k][n].assign(nd(), 0.0f);
                else std::vector<float>().swap(mat[k][n]);
            }
    }

    double DiscLine(int n, uint32_t pos, bool dual) const {
        if (pos >= N[n]) return 0.0;
        if (!dual) return lines[n][pos];
        if (pos < N[n] - 1) return 0.5 * (lines[n][pos] + lines[n][pos + 1]);
        return lines[n][pos] + 0.5 * (lines[n][pos] - lines[n][pos - 1]);
    }
    double DiscDelta(int n, uint32_t pos, bool dual) const {
        if (pos >= N[n]) return 0.0;
        if (!dual) {
            if (pos < N[n] - 1) return DiscLine(n, pos + 1, false) - DiscLine(n, pos, false);
            return DiscLine(n, pos, false) - DiscLine(n, pos - 1, false);
        }
        if (pos > 0) return DiscLine(n, pos, true) - DiscLine(n, pos - 1, true);
        return DiscLine(n, 1, false) - DiscLine(n, 0, false);
    }
    double RawDiscDelta(int n, int pos) const {
        if (pos < 0) return lines[n][0] - lines[n][1];
        if (pos >= (int)N[n] - 1) return lines[n][N[n] - 2] - lines[n][N[n] - 1];
        return lines[n][pos + 1] - lines[n][pos];
    }
    // edge_len[dual][n][i] = DiscDelta(n, i, dual) * gridDelta, filled at build start
    std::vector<double> edge_len[2][3];
    void FillEdgeTables() {
        for (int dl = 0; dl < 2; ++dl)
            for (int n = 0; n < 3; ++n) {
                edge_len[dl][n].resize(N[n]);
                for (uint32_t i = 0; i < N[n]; ++i) edge_len[dl][n][i] = DiscDelta(n, i, dl != 0) * gridDelta;
            }
        // z_shift[z]: the mesh around line z is the mesh around z-1 moved by one
        // line, bit for bit (raw deltas and both edge-length tables over z-3..z+2),
        // so every mesh quantity a cell's EC reads is the same for z and z-1
        FillXCandidates();      // both builds start here
        const uint32_t nz = N[2];
        z_shift.assign(nz, 0);
        if (!row_reuse) return;
        auto same = [](double a, double b) { return memcmp(&a, &b, sizeof a) == 0; };
        for (uint32_t z = 3; z + 3 <= nz; ++z) {
            bool ok = true;
            for (int k = -2; k <= 2 && ok; ++k) {
                const int a = (int)z + k, b = a - 1;
                ok = same(RawDiscDelta(2, a), RawDiscDelta(2, b)) && same(edge_len[0][2][a], edge_len[0][2][b]) &&
                     same(edge_len[1][2][a], edge_len[1][2][b]);
            }
            z_shift[z] = ok;
        }
    }
    // reuse of a row's previous cell where the result is provably the same
    // (EC in ECRow, the Var3 estimate, the coefficients); off = every cell
    // evaluated, for checking
    bool row_reuse = true;
    std::vector<uint8_t> z_shift;
    double EdgeLength(int n, const uint32_t* pos, bool dual = false) const {
        const std::vector<double>& t = edge_len[dual ? 1 : 0][n];
        if (pos[n] < t.size()) return t[pos[n]];
        return DiscDelta(n, pos[n], dual) * gridDelta;
    }
    double NodeWidth(int n, const uint32_t* pos, bool dual) const { return EdgeLength(n, pos, !dual); }
    double NodeWidthI(int n, const int* pos, bool dual) const {
        if (pos[0] < 0 || pos[1] < 0 || pos[2] < 0) return 0.0;
        uint32_t u[3] = {(uint32_t)pos[0], (uint32_t)pos[1], (uint32_t)pos[2]};
        return NodeWidth(n, u, dual);
    }
    double NodeArea(int n, const uint32_t* pos, bool dual) const {
        return NodeWidth((n + 1) % 3, pos, dual) * NodeWidth((n + 2) % 3, pos, dual);
    }
    double NodeAreaI(int n, const int* pos, bool dual) const {
        if (pos[0] < 0 || pos[1] < 0 || pos[2] < 0) return 0.0;
        uint32_t u[3] = {(uint32_t)pos[0], (uint32_t)pos[1], (uint32_t)pos[2]};
        return NodeArea(n, u, dual);
    }
    double EdgeArea(int n, const uint32_t* pos, bool dual = false) const { return NodeArea(n, pos, dual); }
    // GetYeeCoords, FDTD/operator.cpp:174-194
    bool YeeCoords(int ny, const uint32_t* pos, double* c, bool dual) const {
        for (int n = 0; n < 3; ++n) c[n] = DiscLine(n, pos[n], dual);
        c[ny] = DiscLine(ny, pos[ny], !dual);
        if (!dual) return pos[ny] < N[ny] - 1;
        int nP = (ny + 1) % 3, nPP = (ny + 2) % 3;
        return !(pos[nP] >= N[nP] - 1 || pos[nPP] >= N[nPP] - 1);
    }
    uint32_t SnapToMeshLine(int ny, double coord, bool& inside, bool dual) const {
        inside = false;
        if (coord < DiscLine(ny, 0, false)) return 0;
        if (coord > DiscLine(ny, N[ny] - 1, false)) return N[ny] - 1;
        inside = true;
        if (!dual) {
            for (uint32_t n = 0; n < N[ny]; ++n)
                if (coord <= DiscLine(ny, n, true)) return n;
        } else {
            for (uint32_t n = 1; n < N[ny]; ++n)
                if (coord <= DiscLine(ny, n, false)) return n - 1;
        }
        return 0;
    }

    // ---- geometry lookup (CSX->GetPropertyByCoordPriority) ----
    // primitives are kept in descending priority order; the first primitive
    // of a matching kind that contains the point wins
    int PropertyAt(const double* c, int kinds, const std::vector<int>* cand = nullptr) const {
        if (cand) {
            for (int i : *cand) {
                const Primitive& p = prims[i];
                if ((props[p.prop].kind & kinds) && p.inside(c)) return p.prop;
            }
            return -1;
        }
        for (const Primitive& p : prims)
            if ((props[p.prop].kind & kinds) && p.inside(c)) return p.prop;
        return -1;
    }
    // candidate primitives whose bounding box touches [lo,hi] (drawing units)
    std::vector<int> Candidates(const double lo[3], const double hi[3], int kinds) const {
        std::vector<int> r;
        for (size_t i = 0; i < prims.size(); ++i) {
            const Primitive& p = prims[i];
            if (!(props[p.prop].kind & kinds)) continue;
            bool ok = true;
            for (int n = 0; n < 3; ++n)
                if (p.bb1[n] < lo[n] || p.bb0[n] > hi[n]) ok = false;
            if (ok) r.push_back((int)i);
        }
        return r;
    }
    // the x (n = 0) or y (n = 1) window of the rows at line p: one cell either
    // side, widened by its length (quarter-cell and dual sample points stay
    // within one cell of the row)
    void RowWindow(int n, int p, double& lo, double& hi) const {
        lo = lines[n][std::max(0, p - 1)];
        hi = lines[n][std::min((int)N[n] - 1, p + 1)];
        double d = (hi - lo);
        lo -= d; hi += d;
    }
    // xcand[x]: the primitives whose bounding box meets the x window of line
    // x, in ascending index (priority) order; filled at build start, cleared
    // when the primitives or the mesh change
    std::vector<std::vector<int>> xcand;
    void FillXCandidates() {
        xcand.assign(N[0], {});
        for (uint32_t x = 0; x < N[0]; ++x) {
            double lo, hi;
            RowWindow(0, (int)x, lo, hi);
            for (size_t i = 0; i < prims.size(); ++i)
                if (!(prims[i].bb1[0] < lo || prims[i].bb0[0] > hi)) xcand[x].push_back((int)i);
        }
    }
    std::vector<int> RowCandidates(int x, int y, int kinds) const {
        double lo[3], hi[3];
        int pos[2] = {x, y};
        for (int n = 0; n < 2; ++n) RowWindow(n, pos[n], lo[n], hi[n]);
        lo[2] = -std::numeric_limits<double>::infinity();
        hi[2] = std::numeric_limits<double>::infinity();
        if (xcand.size() != N[0]) return Candidates(lo, hi, kinds);
        // the same test as Candidates over the primitives that pass its x test
        std::vector<int> r;
        for (int i : xcand[x]) {
            const Primitive& p = prims[i];
            if (!(props[p.prop].kind & kinds)) continue;
            bool ok = true;
            for (int n = 0; n < 3; ++n)
                if (p.bb1[n] < lo[n] || p.bb0[n] > hi[n]) ok = false;
            if (ok) r.push_back(i);
        }
        return r;
    }

    // GetMaterial (FDTD/operator.cpp:1251-1298)
    double GetMaterial(const double* c, int type, const std::vector<int>& cand) const {
        int p = PropertyAt(c, P_MATERIAL, &cand);
        if (p >= 0) {
            const Property& q = props[p];
            switch (type) {
            case 0: return q.eps;
            case 1: return q.kappa;
            case 2: return q.mue;
            default: return q.sigma;
            }
        }
        switch (type) {
        case 0: return 1.0;
        case 1: return 0.0;
        case 2: return 1.0;
        default: return 0.0;
        }
    }

    // two GetMaterial values of one sample point with a single lookup
    void Material2(const double* c, const std::vector<int>& cand, int ta, int tb, double& a, double& b) const {
        const int p = PropertyAt(c, P_MATERIAL, &cand);
        auto val = [&](int type) -> double {
            if (p >= 0) {
                const Property& q = props[p];
                switch (type) {
                case 0: return q.eps;
                case 1: return q.kappa;
                case 2: return q.mue;
                default: return q.sigma;
                }
            }
            return (type == 0 || type == 2) ? 1.0 : 0.0;
        };
        a = val(ta);
        b = val(tb);
    }

    // AverageMatQuarterCell (FDTD/operator.cpp:1406-1518)
    void AverageMatQuarterCell(int ny, const uint32_t* pos, double* EffMat, const std::vector<int>& cand) const {
        const int n = ny, nP = (n + 1) % 3, nPP = (n + 2) % 3;
        double coord[3] = {lines[0][pos[0]], lines[1][pos[1]], lines[2][pos[2]]};
        double shiftCoord[3];
        const double delta = RawDiscDelta(n, pos[n]);
        const double deltaP = RawDiscDelta(nP, pos[nP]);
        const double deltaPP = RawDiscDelta(nPP, pos[nPP]);
        const double delta_M = RawDiscDelta(n, (int)pos[n] - 1);
        const double deltaP_M = RawDiscDelta(nP, (int)pos[nP] - 1);
        const double deltaPP_M = RawDiscDelta(nPP, (int)pos[nPP] - 1);
        int loc_pos[3] = {(int)pos[0], (int)pos[1], (int)pos[2]};
        double A_n, area = 0;

        shiftCoord[n] = coord[n] + delta * 0.5;
        shiftCoord[nP] = coord[nP] + deltaP * 0.25;
        shiftCoord[nPP] = coord[nPP] + deltaPP * 0.25;
        A_n = NodeAreaI(ny, loc_pos, true);
        {
            double e, k;
            Material2(shiftCoord, cand, 0, 1, e, k);
            EffMat[0] = e * A_n;
            EffMat[1] = k * A_n;
        }
        area += A_n;

        shiftCoord[n] = coord[n] + delta * 0.5;
        shiftCoord[nP] = coord[nP] - deltaP_M * 0.25;
        shiftCoord[nPP] = coord[nPP] + deltaPP * 0.25;
        --loc_pos[nP];
        A_n = NodeAreaI(ny, loc_pos, true);
        {
            double e, k;
            Material2(shiftCoord, cand, 0, 1, e, k);
            EffMat[0] += e * A_n;
            EffMat[1] += k * A_n;
        }
        area += A_n;

        shiftCoord[n] = coord[n] + delta * 0.5;
        shiftCoord[nP] = coord[nP] + deltaP * 0.25;
        shiftCoord[nPP] = coord[nPP] - deltaPP_M * 0.25;
        ++loc_pos[nP];
        --loc_pos[nPP];
        A_n = NodeAreaI(ny, loc_pos, true);
        {
            double e, k;
            Material2(shiftCoord, cand, 0, 1, e, k);
            EffMat[0] += e * A_n;
            EffMat[1] += k * A_n;
        }
        area += A_n;

        shiftCoord[n] = coord[n] + delta * 0.5;
        shiftCoord[nP] = coord[nP] - deltaP_M * 0.25;
        shiftCoord[nPP] = coord[nPP] - deltaPP_M * 0.25;
        --loc_pos[nP];
        A_n = NodeAreaI(ny, loc_pos, true);
        {
            double e, k;
            Material2(shiftCoord, cand, 0, 1, e, k);
            EffMat[0] += e * A_n;
            EffMat[1] += k * A_n;
        }
        area += A_n;

        EffMat[0] *= EPS0 / area;
        EffMat[1] /= area;

        loc_pos[0] = pos[0]; loc_pos[1] = pos[1]; loc_pos[2] = pos[2];
        double length = 0;
        shiftCoord[n] = coord[n] - delta_M * 0.25;
        shiftCoord[nP] = coord[nP] + deltaP * 0.5;
        shiftCoord[nPP] = coord[nPP] + deltaPP * 0.5;
        --loc_pos[n];
        double delta_ny = NodeWidthI(n, loc_pos, true);
        double mue, sigma;
        Material2(shiftCoord, cand, 2, 3, mue, sigma);
        EffMat[2] = delta_ny / mue;
        if (sigma) EffMat[3] = delta_ny / sigma;
        else EffMat[3] = 0;
        length = delta_ny;

        shiftCoord[n] = coord[n] + delta * 0.25;
        shiftCoord[nP] = coord[nP] + deltaP * 0.5;
        shiftCoord[nPP] = coord[nPP] + deltaPP * 0.5;
        ++loc_pos[n];
        delta_ny = NodeWidthI(n, loc_pos, true);
        Material2(shiftCoord, cand, 2, 3, mue, sigma);
        EffMat[2] += delta_ny / mue;
        if (sigma) EffMat[3] += delta_ny / sigma;
        else EffMat[3] = 0;
        length += delta_ny;

        EffMat[2] = length * MUE0 / EffMat[2];
        if (EffMat[3]) EffMat[3] = length / EffMat[3];
    }

    template <typename F>
    F rnd(double v) const { return (F)v; }

    int threads() const {
        int t = nthreads > 0 ? nthreads : (int)std::thread::hardware_concurrency();
        return std::max(1, std::min(t, 64));
    }

    // the build's workers (PoolScope); none are kept between calls, so an
    // idle operator holds no threads
    mutable std::unique_ptr<WorkerPool> pool;
    mutable pid_t pool_pid = 0;
    mutable int pool_depth = 0;
    template <typename Fn>
    void parallel_x(uint32_t nx, Fn fn) const {
        const int T = threads();
        if (T <= 1 || nx <= 1 || WorkerPool::inside()) {
            for (uint32_t x = 0; x < nx; ++x) fn(x);
            return;
        }
        if (pool_depth == 0) {
            std::atomic<uint32_t> next(0);
            std::vector<std::thread> th;
            for (int t = 0; t < T; ++t)
                th.emplace_back([&]() {
                    for (uint32_t x; (x = next.fetch_add(1)) < nx;) fn(x);
                });
            for (auto& t : th) t.join();
            return;
        }
        // a forked child has none of the parent's workers: leave that pool alone
        if (pool && pool_pid != getpid()) (void)pool.release();
        if (!pool || (int)pool->workers() != T - 1) {
            pool_pid = getpid();
            pool.reset();
            pool.reset(new WorkerPool(T - 1));
        }
        const std::function<void(uint32_t)> f = [&](uint32_t x) { fn(x); };
        pool->run(nx, f);
    }

    template <typename F>
    int build(uint32_t nr_ts);
};

namespace {

// ---- coefficient helpers with F rounding at the reference's storage points ----
// Along one row of cells (x, y fixed, z ascending): whether
// AverageMatQuarterCell of cell z gives what it gave for cell z-1.  Its
// sample points for cell z lie in [line z-1, line z+1] and share x and y
// with those of cell z-1, and a box's inside test along the row changes only
// at its z faces: with no face of a candidate box in [line z-2, line z+1] and
// the mesh around z a bit-exact shift of the mesh around z-1 (z_shift), the
// results are the same.  Spheres are not interval-shaped in z: rows that can
// see one evaluate every cell.
struct ZRunReuse {
    const wop_operator* op;
    std::vector<double> faces;
    size_t fp = 0;
    bool on;
    ZRunReuse(const wop_operator* o, const std::vector<int>& cand, bool enable) : op(o), on(enable && !o->z_shift.empty()) {
        for (int c : cand) {
            const Primitive& p = op->prims[c];
            if (p.shape != 0) on = false;
            faces.push_back(p.a[2]);
            faces.push_back(p.b[2]);
        }
        std::sort(faces.begin(), faces.end());
    }
    // z ascending over the calls
    bool same(uint32_t z) {
        if (!on || !op->z_shift[z]) return false;
        const double lo = op->lines[2][z - 2], hi = op->lines[2][z + 1];
        while (fp < faces.size() && faces[fp] < lo) ++fp;
        return fp == faces.size() || faces[fp] > hi;
    }
};

template <typename F>
struct Builder {
    wop_operator* op;
    std::vector<F> (&EC)[4][3];   // C,G,L,R per direction
    std::vector<F> (&co)[4][3];   // vv, vi, ii, iv per direction
    // same_prev[i] = 1: ECRow copied cell i's EC (all directions) from the
    // cell below it in z, whose EC is then bitwise the same
    std::vector<uint8_t> same_prev;

    explicit Builder(wop_operator* o) : op(o), EC(o->store<F>().EC), co(o->store<F>().co) {}

    F& VV(int n, size_t i) { return co[0][n][i]; }
    F& VI(int n, size_t i) { return co[1][n][i]; }
    F& II(int n, size_t i) { return co[2][n][i]; }
    F& IV(int n, size_t i) { return co[3][n][i]; }

    // Calc_ECOperatorPos, FDTD/operator.cpp:943-971
    void ECOperatorPos(int n, size_t i) {
        const double dT = op->dT;
        double C = EC[0][n][i], G = EC[1][n][i];
        if (C > 0) {
            VV(n, i) = (F)((1.0 - dT * G / 2.0 / C) / (1.0 + dT * G / 2.0 / C));
            VI(n, i) = (F)((dT / C) / (1.0 + dT * G / 2.0 / C));
        } else {
            VV(n, i) = 0;
            VI(n, i) = 0;
        }
        double L = EC[2][n][i], R = EC[3][n][i];
        if (L > 0) {
            II(n, i) = (F)((1.0 - dT * R / 2.0 / L) / (1.0 + dT * R / 2.0 / L));
            IV(n, i) = (F)((dT / L) / (1.0 + dT * R / 2.0 / L));
        } else {
            II(n, i) = 0;
            IV(n, i) = 0;
        }
    }

    // Calc_EC_Range + Calc_ECPos (FDTD/operator.cpp:1784-1809,1176-1224)
    void CalcEC() {
        const size_t nd = op->nd();
        for (int k = 0; k < 4; ++k)
            for (int n = 0; n < 3; ++n) EC[k][n].assign(nd, (F)0);
        same_prev.assign(nd, 0);
        op->parallel_x(op->N[0], [&](uint32_t x) {
            for (uint32_t y = 0; y < op->N[1]; ++y) ECRow(x, y);
        });
    }

    // EC of the cells of row (x, y)
    void ECRow(uint32_t x, uint32_t y) {
            uint32_t pos[3];
            pos[0] = x;
            pos[1] = y;
            double EffMat[4];
            {
                std::vector<int> cand = op->RowCandidates((int)pos[0], (int)pos[1], P_MATERIAL);
                // a cell's EC is a function of its AverageMatQuarterCell results and
                // of the mesh around it: where both repeat, copy cell z-1's EC
                ZRunReuse run(op, cand, !same_prev.empty());
                for (pos[2] = 0; pos[2] < op->N[2]; ++pos[2]) {
                    const size_t i = op->idx(pos[0], pos[1], pos[2]);
                    const size_t li = ((size_t)pos[0] * op->N[1] + pos[1]) * op->N[2] + pos[2];
                    if (run.same(pos[2])) {
                        const size_t ip = op->idx(pos[0], pos[1], pos[2] - 1);
                        for (int k = 0; k < 4; ++k)
                            for (int n = 0; n < 3; ++n) EC[k][n][i] = EC[k][n][ip];
                        for (int k = 0; k < 3; ++k)
                            if (op->store_mat[k])
                                for (int n = 0; n < 3; ++n) op->mat[k][n][li] = op->mat[k][n][li - 1];
                        same_prev[i] = 1;
                        continue;
                    }
                    if (!same_prev.empty()) same_prev[i] = 0;
                    for (int n = 0; n < 3; ++n) {
                        op->AverageMatQuarterCell(n, pos, EffMat, cand);
                        for (int k = 0; k < 3; ++k)
                            if (op->store_mat[k]) op->mat[k][n][li] = (float)EffMat[k];
                        double delta = op->EdgeLength(n, pos);
                        double area = op->EdgeArea(n, pos);
                        double e0, e1, e2, e3;
                        if (delta) { e0 = EffMat[0] * area / delta; e1 = EffMat[1] * area / delta; }
                        else { e0 = 0; e1 = 0; }
                        delta = op->EdgeLength(n, pos, true);
                        area = op->EdgeArea(n, pos, true);
                        if (delta) { e2 = EffMat[2] * area / delta; e3 = EffMat[3] * area / delta; }
                        else { e2 = 0; e3 = 0; }
                        EC[0][n][i] = (F)e0; EC[1][n][i] = (F)e1; EC[2][n][i] = (F)e2; EC[3][n][i] = (F)e3;
                    }
                }
            }
    }

    void ECPlane(uint32_t x) {
        op->parallel_x(op->N[1], [&](uint32_t y) { ECRow(x, y); });
    }

    // CalcTimestep_Var3 (FDTD/operator.cpp:1897-1971), AdrOp with
    // SetReflection2Cell: a shift of -1 at index 0 or +1 at the last index
    // stays on the same cell (tools/AdrOp.cpp:183-220)
    double CalcTimestep() {
        const uint32_t* N = op->N;
        double dT = 1e200;
        for (int n = 0; n < 3; ++n) {
            // per-z-plane minima computed in parallel, combined in the reference's loop order
            std::vector<double> zmin(N[2], 1e200);
            op->parallel_x(N[2], [&](uint32_t z) {
                uint32_t pos[3];
                pos[2] = z;
                double best = 1e200;
                for (pos[1] = 0; pos[1] < N[1]; ++pos[1]) {
                    for (pos[0] = 0; pos[0] < N[0]; ++pos[0]) {
                        const double newT = Var3Cell(n, pos);
                        if ((newT < best) && (newT > 0.0)) best = newT;
                    }
                }
                zmin[z] = best;
            });
            for (uint32_t z = 0; z < N[2]; ++z)
                if (zmin[z] < dT && zmin[z] > 0.0) dT = zmin[z];
        }
        return dT;
    }

    // smallest positive Var3 timestep of the cells of plane x (EC of x-1..x+1 present)
    double TimestepPlane(uint32_t x) {
        const uint32_t* N = op->N;
        std::vector<double> ymin(N[1], 1e200);
        op->parallel_x(N[1], [&](uint32_t y) {
            uint32_t pos[3] = {x, y, 0};
            // Var3Cell(n, z) reads EC within one cell of (x, y, z) on each axis
            // (x and y reflected at the ends, z at most one away).  Where the 3x3
            // columns around the row all have EC(z-2) = EC(z-1) = EC(z) = EC(z+1),
            // the estimate at z is the one at z-1, already in the minimum.
            std::vector<uint8_t> col(N[2], 0);
            if (!same_prev.empty()) {
                for (uint32_t z = 0; z < N[2]; ++z) col[z] = 1;
                auto refl = [](int q, int n) { return q < 0 ? -q - 1 : (q > n - 1 ? 2 * (n - 1) - q + 1 : q); };
                for (int dx = -1; dx <= 1; ++dx)
                    for (int dy = -1; dy <= 1; ++dy) {
                        const int qx = refl((int)x + dx, (int)N[0]), qy = refl((int)y + dy, (int)N[1]);
                        const size_t i0 = op->idx(qx, qy, 0);
                        for (uint32_t z = 0; z < N[2]; ++z) col[z] &= same_prev[i0 + z];
                    }
            }
            double best = 1e200;
            for (pos[2] = 0; pos[2] < N[2]; ++pos[2]) {
                const uint32_t z = pos[2];
                if (z >= 1 && z + 1 < N[2] && col[z - 1] && col[z] && col[z + 1]) continue;
                for (int n = 0; n < 3; ++n) {
                    const double newT = Var3Cell(n, pos);
                    if ((newT < best) && (newT > 0.0)) best = newT;
                }
            }
            ymin[y] = best;
        });
        double dT = 1e200;
        for (uint32_t y = 0; y < N[1]; ++y)
            if (ymin[y] < dT && ymin[y] > 0.0) dT = ymin[y];
        return dT;
    }

    // Var3 timestep estimate of direction n at one cell
    double Var3Cell(int n, const uint32_t* pos) {
        const uint32_t* N = op->N;
        auto sh = [&](const uint32_t* p, int dir, int step) -> size_t {
            int q[3] = {(int)p[0], (int)p[1], (int)p[2]};
            q[dir] += step;
            if (q[dir] < 0) q[dir] = -q[dir] - 1;
            if (q[dir] > (int)N[dir] - 1) q[dir] = 2 * ((int)N[dir] - 1) - q[dir] + 1;
            return op->idx(q[0], q[1], q[2]);
        };
        auto sh2 = [&](const uint32_t* p, int d1, int s1, int d2, int s2) -> size_t {
            // GetShiftedPos after a Shift: both offsets applied, then reflected per axis
            int q[3] = {(int)p[0], (int)p[1], (int)p[2]};
            q[d1] += s1;
            q[d2] += s2;
            for (int d = 0; d < 3; ++d) {
                if (q[d] < 0) q[d] = -q[d] - 1;
                if (q[d] > (int)N[d] - 1) q[d] = 2 * ((int)N[d] - 1) - q[d] + 1;
            }
            return op->idx(q[0], q[1], q[2]);
        };
        const int nP = (n + 1) % 3, nPP = (n + 2) % 3;
        const std::vector<F>& Cn = EC[0][n];
        const std::vector<F>& CnP = EC[0][nP];
        const std::vector<F>& CnPP = EC[0][nPP];
        const std::vector<F>& LnP = EC[2][nP];
        const std::vector<F>& LnPP = EC[2][nPP];
        // EC_* are FDTD_FLOAT: each 1/(L*C) term and the pair sums are
        // evaluated in F, then accumulated in double (operator.cpp:1920-1951)
        const F one = (F)1;
        size_t ipos = op->idx(pos[0], pos[1], pos[2]);
        double wqp;
        wqp = (double)(one / (LnPP[ipos] * Cn[sh(pos, nP, 1)]) + one / (LnPP[ipos] * Cn[ipos]));
        wqp += (double)(one / (LnP[ipos] * Cn[sh(pos, nPP, 1)]) + one / (LnP[ipos] * Cn[ipos]));
        // Shift(nP,-1): GetShiftedPos(nP,1) from there is the original cell
        size_t ip1 = sh(pos, nP, -1);
        wqp += (double)(one / (LnPP[ip1] * Cn[sh2(pos, nP, -1, nP, 1)]) + one / (LnPP[ip1] * Cn[ip1]));
        // Shift(nPP,-1) keeps the nP shift of the previous Shift call
        size_t ip2 = sh2(pos, nP, -1, nPP, -1);
        wqp += (double)(one / (LnP[ip2] * Cn[sh2(pos, nP, -1, nPP, 0)]) + one / (LnP[ip2] * Cn[ip2]));

        double wt_4[4];
        wt_4[0] = (double)(one / (LnPP[ipos] * CnP[ipos]));
        wt_4[1] = (double)(one / (LnPP[sh(pos, nP, -1)] * CnP[ipos]));
        wt_4[2] = (double)(one / (LnP[ipos] * CnPP[ipos]));
        wt_4[3] = (double)(one / (LnP[sh(pos, nPP, -1)] * CnPP[ipos]));
        double wt1 = wt_4[0] + wt_4[1] + wt_4[2] + wt_4[3] - 2 * std::min(std::min(wt_4[0], wt_4[1]), std::min(wt_4[2], wt_4[3]));

        size_t in1 = sh(pos, n, 1);
        wt_4[0] = (double)(one / (LnPP[ipos] * CnP[in1]));
        wt_4[1] = (double)(one / (LnPP[sh(pos, nP, -1)] * CnP[in1]));
        wt_4[2] = (double)(one / (LnP[ipos] * CnPP[in1]));
        wt_4[3] = (double)(one / (LnP[sh(pos, nPP, -1)] * CnPP[in1]));
        double wt2 = wt_4[0] + wt_4[1] + wt_4[2] + wt_4[3] - 2 * std::min(std::min(wt_4[0], wt_4[1]), std::min(wt_4[2], wt_4[3]));

        double w_total = wqp + wt1 + wt2;
        return 2 / sqrt(w_total);
    }

    // cells of plane only_x (all cells when only_x < 0): loop bounds per axis
    void AxisRange(int only_x, uint32_t lo[3], uint32_t hi[3]) const {
        for (int d = 0; d < 3; ++d) { lo[d] = 0; hi[d] = op->N[d]; }
        if (only_x >= 0) { lo[0] = (uint32_t)only_x; hi[0] = (uint32_t)only_x + 1; }
    }

    void ApplyElectricBC(int only_x = -1) {
        bool dirs[6];
        for (int k = 0; k < 6; ++k) dirs[k] = op->BC[k] != -1;
        const uint32_t* N = op->N;
        uint32_t lo[3], hi[3];
        AxisRange(only_x, lo, hi);
        uint32_t pos[3];
        for (int n = 0; n < 3; ++n) {
            int nP = (n + 1) % 3, nPP = (n + 2) % 3;
            const bool in_lo = n != 0 || only_x < 0 || only_x == 0;
            const bool in_hi = n != 0 || only_x < 0 || only_x == (int)N[0] - 1;
            if (!(dirs[2 * n] && in_lo) && !(dirs[2 * n + 1] && in_hi)) continue;   // nothing on this plane
            for (pos[nP] = lo[nP]; pos[nP] < hi[nP]; ++pos[nP])
                for (pos[nPP] = lo[nPP]; pos[nPP] < hi[nPP]; ++pos[nPP]) {
                    if (dirs[2 * n] && in_lo) {
                        pos[n] = 0;
                        size_t i = op->idx(pos[0], pos[1], pos[2]);
                        VV(nP, i) = 0; VI(nP, i) = 0; VV(nPP, i) = 0; VI(nPP, i) = 0;
                    }
                    if (dirs[2 * n + 1] && in_hi) {
                        pos[n] = N[n] - 1;
                        size_t i = op->idx(pos[0], pos[1], pos[2]);
                        VV(n, i) = 0; VI(n, i) = 0;
                        VV(nP, i) = 0; VI(nP, i) = 0; VV(nPP, i) = 0; VI(nPP, i) = 0;
                    }
                }
        }
    }

    void ApplyMagneticBC(int only_x = -1) {
        bool dirs[6];
        for (int k = 0; k < 6; ++k) dirs[k] = op->BC[k] == 1;
        const uint32_t* N = op->N;
        uint32_t lo[3], hi[3];
        AxisRange(only_x, lo, hi);
        uint32_t pos[3];
        for (int n = 0; n < 3; ++n) {
            int nP = (n + 1) % 3, nPP = (n + 2) % 3;
            const bool at = n != 0 || only_x < 0;
            if (!at && only_x != 0 && only_x != (int)N[0] - 2 && only_x != (int)N[0] - 1) continue;   // nothing on this plane
            for (pos[nP] = lo[nP]; pos[nP] < hi[nP]; ++pos[nP])
                for (pos[nPP] = lo[nPP]; pos[nPP] < hi[nPP]; ++pos[nPP]) {
                    if (dirs[2 * n] && (at || only_x == 0)) {
                        pos[n] = 0;
                        size_t i = op->idx(pos[0], pos[1], pos[2]);
                        II(n, i) = 0; IV(n, i) = 0; II(nP, i) = 0; IV(nP, i) = 0; II(nPP, i) = 0; IV(nPP, i) = 0;
                    }
                    if (dirs[2 * n + 1] && (at || only_x == (int)N[0] - 2)) {
                        pos[n] = N[n] - 2;
                        size_t i = op->idx(pos[0], pos[1], pos[2]);
                        II(nP, i) = 0; IV(nP, i) = 0; II(nPP, i) = 0; IV(nPP, i) = 0;
                    }
                    if (at || only_x == (int)N[0] - 1) {
                        pos[n] = N[n] - 1;
                        size_t i = op->idx(pos[0], pos[1], pos[2]);
                        II(n, i) = 0; IV(n, i) = 0; II(nP, i) = 0; IV(nP, i) = 0; II(nPP, i) = 0; IV(nPP, i) = 0;
                    }
                }
        }
    }

    // CalcPEC_Range (FDTD/operator.cpp:1986-2020): edges whose Yee point lies in metal
    void CalcPEC() {
        op->parallel_x(op->N[0], [&](uint32_t x) {
            for (uint32_t y = 0; y < op->N[1]; ++y) PECRow(x, y);
        });
    }

    void PECPlane(uint32_t x) {
        op->parallel_x(op->N[1], [&](uint32_t y) { PECRow(x, y); });
    }

    void PECRow(uint32_t x, uint32_t y) {
        uint32_t pos[3] = {x, y, 0};
        double coord[3];
        std::vector<int> cand = op->RowCandidates((int)pos[0], (int)pos[1], P_MATERIAL | P_METAL);
        bool any_metal = false;
        for (int c : cand) if (op->props[op->prims[c].prop].kind & P_METAL) any_metal = true;
        if (!any_metal) return;
        for (pos[2] = 0; pos[2] < op->N[2]; ++pos[2])
            for (int n = 0; n < 3; ++n) {
                op->YeeCoords(n, pos, coord, false);
                int p = op->PropertyAt(coord, P_MATERIAL | P_METAL, &cand);
                if (p >= 0 && (op->props[p].kind & P_METAL)) {
                    size_t i = op->idx(pos[0], pos[1], pos[2]);
                    VV(n, i) = 0;
                    VI(n, i) = 0;
                }
            }
    }

    // Calc_LumpedElements (FDTD/operator.cpp:1535-1712), parallel RC boxes only
    int LumpedElements(int only_x = -1) {
        for (const Primitive& pr : op->prims) {
            const Property& P = op->props[pr.prop];
            if (!(P.kind & P_LUMPED) || pr.shape != 0) continue;
            // only parallel RC here (Operator::IsLEparRC, FDTD/operator.cpp:1714-1724)
            if (!(P.le_type == 0 && !(P.L > 0.0))) continue;
            double C = P.C;
            if (C <= 0) C = NAN;
            double R = P.R;
            if (R < 0) R = NAN;
            if (std::isnan(R) && std::isnan(C)) continue;
            int ny = P.ny;
            if (ny < 0 || ny > 2) continue;
            int nyP = (ny + 1) % 3, nyPP = (ny + 2) % 3;
            uint32_t s[3], e[3];
            bool in;
            for (int n = 0; n < 3; ++n) {
                s[n] = op->SnapToMeshLine(n, pr.a[n], in, false);
                e[n] = op->SnapToMeshLine(n, pr.b[n], in, false);
            }
            if (s[ny] == e[ny]) continue;
            uint32_t pos[3];
            double unitGC = 0;
            for (pos[ny] = s[ny]; pos[ny] < e[ny]; ++pos[ny]) {
                double plane = 0;
                for (pos[nyP] = s[nyP]; pos[nyP] <= e[nyP]; ++pos[nyP])
                    for (pos[nyPP] = s[nyPP]; pos[nyPP] <= e[nyPP]; ++pos[nyPP])
                        plane += op->EdgeArea(ny, pos) / op->EdgeLength(ny, pos);
                unitGC += 1 / plane;
            }
            unitGC = 1 / unitGC;
            double kappa = 0, epsilon = 0;
            if (R > 0) kappa = 1 / R / unitGC;
            if (C > 0) {
                epsilon = C / unitGC;
                if (epsilon < EPS0) C = 0;
            }
            for (pos[ny] = s[ny]; pos[ny] < e[ny]; ++pos[ny])
                for (pos[nyP] = s[nyP]; pos[nyP] <= e[nyP]; ++pos[nyP])
                    for (pos[nyPP] = s[nyPP]; pos[nyPP] <= e[nyPP]; ++pos[nyPP]) {
                        if (only_x >= 0 && pos[0] != (uint32_t)only_x) continue;
                        size_t i = op->idx(pos[0], pos[1], pos[2]);
                        if (C > 0) EC[0][ny][i] = (F)(epsilon * op->EdgeArea(ny, pos) / op->EdgeLength(ny, pos));
                        if (R > 0) EC[1][ny][i] = (F)(kappa * op->EdgeArea(ny, pos) / op->EdgeLength(ny, pos));
                        if (R == 0) { VV(ny, i) = 0; VI(ny, i) = 0; }
                        else ECOperatorPos(ny, i);
                    }
            if (P.caps) {
                for (pos[nyP] = s[nyP]; pos[nyP] <= e[nyP]; ++pos[nyP])
                    for (pos[nyPP] = s[nyPP]; pos[nyPP] <= e[nyPP]; ++pos[nyPP]) {
                        for (int side = 0; side < 2; ++side) {
                            pos[ny] = side ? e[ny] : s[ny];
                            if (only_x >= 0 && pos[0] != (uint32_t)only_x) continue;
                            size_t i = op->idx(pos[0], pos[1], pos[2]);
                            if (pos[nyP] < e[nyP]) { VV(nyP, i) = 0; VI(nyP, i) = 0; }
                            if (pos[nyPP] < e[nyPP]) { VV(nyPP, i) = 0; VI(nyPP, i) = 0; }
                        }
                    }
            }
        }
        return 0;
    }

    // Operator_Ext_Excitation::BuildExtension (operator_ext_excitation.cpp:101-222), box primitives
    // only_x >= 0: the entries of plane only_x, appended (the caller clears
    // the lists and restores the z, y, x order afterwards)
    void Excitation(int only_x = -1) {
        if (only_x < 0) {
            op->exc_v.clear();
            op->exc_i.clear();
        }
        const uint32_t* N = op->N;
        const double dT = op->dT;
        bool any = false;
        for (auto& p : op->prims) if (op->props[p.prop].kind & P_EXCITATION) any = true;
        if (!any) return;
        uint32_t lo[3], hi[3];
        AxisRange(only_x, lo, hi);
        // only excitation primitives can match (scanned in priority order);
        // for one plane, only those whose bounding box comes within a cell
        // of the plane's sample points
        const double inf = std::numeric_limits<double>::infinity();
        double blo[3] = {-inf, -inf, -inf}, bhi[3] = {inf, inf, inf};
        if (only_x >= 0) {
            const auto& L = op->lines[0];
            blo[0] = L[std::max(0, only_x - 1)];
            bhi[0] = L[std::min((int)N[0] - 1, only_x + 1)];
            const double d = bhi[0] - blo[0];
            blo[0] -= d;
            bhi[0] += d;
        }
        const std::vector<int> cand = op->Candidates(blo, bhi, P_EXCITATION);
        if (cand.empty()) return;
        uint32_t pos[3];
        double c[3];
        for (pos[2] = 0; pos[2] < N[2]; ++pos[2])
            for (pos[1] = 0; pos[1] < N[1]; ++pos[1])
                for (pos[0] = lo[0]; pos[0] < hi[0]; ++pos[0]) {
                    for (int n = 0; n < 3; ++n) {
                        if (!op->YeeCoords(n, pos, c, false)) continue;
                        int p = op->PropertyAt(c, P_EXCITATION, &cand);
                        if (p < 0) continue;
                        const Property& P = op->props[p];
                        if (P.exc_val[n] != 0 && (P.exc_type == 0 || P.exc_type == 1)) {
                            double amp = P.exc_val[n] * op->EdgeLength(n, pos);
                            if (amp != 0) {
                                wop_operator::Exc e;
                                for (int k = 0; k < 3; ++k) e.pos[k] = pos[k];
                                e.dir = n;
                                e.amp = (double)(F)amp;
                                e.delay = (uint32_t)(P.delay / dT);
                                op->exc_v.push_back(e);
                            }
                            if (P.exc_type == 1) {
                                size_t i = op->idx(pos[0], pos[1], pos[2]);
                                VV(n, i) = 0; VI(n, i) = 0;
                            }
                        }
                    }
                    for (int n = 0; n < 3; ++n) {
                        if (pos[0] >= N[0] - 1 || pos[1] >= N[1] - 1 || pos[2] >= N[2] - 1) continue;
                        if (!op->YeeCoords(n, pos, c, true)) continue;
                        int p = op->PropertyAt(c, P_EXCITATION, &cand);
                        if (p < 0) continue;
                        const Property& P = op->props[p];
                        if (P.exc_val[n] != 0 && (P.exc_type == 2 || P.exc_type == 3)) {
                            double amp = P.exc_val[n] * op->EdgeLength(n, pos, true);
                            if (amp != 0) {
                                wop_operator::Exc e;
                                for (int k = 0; k < 3; ++k) e.pos[k] = pos[k];
                                e.dir = n;
                                e.amp = (double)(F)amp;
                                e.delay = (uint32_t)(P.delay / dT);
                                op->exc_i.push_back(e);
                            }
                            if (P.exc_type == 3) {
                                size_t i = op->idx(pos[0], pos[1], pos[2]);
                                II(n, i) = 0; IV(n, i) = 0;
                            }
                        }
                    }
                }
    }

    // default grading function of Operator_Ext_UPML (operator_ext_upml.cpp:31):
    // -log(1e-6)*log(2.5)/(2*dl*Z*(pow(2.5,W/dl)-1)) * pow(2.5, D/dl)
    static double Grading(double D, double dl, double W, double Z, double) {
        return -log(1e-6) * log(2.5) / (2 * dl * Z * (pow(2.5, W / dl) - 1)) * pow(2.5, D / dl);
    }

    // GradingKappa's per-axis factors depend only on pos[n] and whether n is
    // the edge direction: tabulated once per build as [n][n == ny][pos[n]]
    std::vector<double> gk_v[3][2], gk_i[3][2];
    void GradingTables(double Zm) {
        for (int n = 0; n < 3; ++n) {
            const int other = (n + 1) % 3;
            for (int isny = 0; isny < 2; ++isny) {
                gk_v[n][isny].assign(op->N[n], 0.0);
                gk_i[n][isny].assign(op->N[n], 0.0);
                for (uint32_t q = 0; q < op->N[n]; ++q) {
                    uint32_t pos[3] = {0, 0, 0};
                    pos[n] = q;
                    double kv[3], ki[3];
                    GradingKappa(isny ? n : other, pos, Zm, kv, ki);
                    gk_v[n][isny][q] = kv[n];
                    gk_i[n][isny][q] = ki[n];
                }
            }
        }
    }
    void GradingKappaTab(int ny, const uint32_t* pos, double* kv, double* ki) const {
        for (int n = 0; n < 3; ++n) {
            const int isny = n == ny;
            kv[n] = gk_v[n][isny][pos[n]];
            ki[n] = gk_i[n][isny][pos[n]];
        }
    }

    // Operator_Ext_UPML::CalcGradingKappa (operator_ext_upml.cpp:289-363)
    void GradingKappa(int ny, const uint32_t* pos, double Zm, double* kv, double* ki) {
        const int* BC = op->BC;
        const uint32_t* S = op->pml_size;
        for (int n = 0; n < 3; ++n) {
            const uint32_t Nn = op->N[n];
            if (pos[n] <= S[2 * n] && BC[2 * n] == 3) {
                double width = (op->lines[n][S[2 * n]] - op->lines[n][0]) * op->gridDelta;
                double depth = width - (op->lines[n][pos[n]] - op->lines[n][0]) * op->gridDelta;
                if (n == ny) depth -= op->EdgeLength(n, pos) / 2;
                double v1 = width / S[2 * n], v4 = (double)S[2 * n];
                kv[n] = depth > 0 ? Grading(depth, v1, width, Zm, v4) : 0;
                if (n == ny) depth += op->EdgeLength(n, pos) / 2;
                if (n != ny) depth -= op->EdgeLength(n, pos) / 2;
                if (depth < 0) depth = 0;
                ki[n] = depth > 0 ? Grading(depth, v1, width, Zm, v4) : 0;
            } else if (pos[n] >= Nn - 1 - S[2 * n + 1] && BC[2 * n + 1] == 3) {
                double width = (op->lines[n][Nn - 1] - op->lines[n][Nn - S[2 * n + 1] - 1]) * op->gridDelta;
                double depth = width - (op->lines[n][Nn - 1] - op->lines[n][pos[n]]) * op->gridDelta;
                if (n == ny) depth += op->EdgeLength(n, pos) / 2;
                // the reference passes the lower-side size here (operator_ext_upml.cpp:338)
                double v1 = width / (S[2 * n]), v4 = (double)S[2 * n];
                kv[n] = depth > 0 ? Grading(depth, v1, width, Zm, v4) : 0;
                if (n == ny) depth -= op->EdgeLength(n, pos) / 2;
                if (n != ny) depth += op->EdgeLength(n, pos) / 2;
                if (depth > width) depth = 0;
                ki[n] = depth > 0 ? Grading(depth, v1, width, Zm, v4) : 0;
            } else {
                kv[n] = 0;
                ki[n] = 0;
            }
        }
    }

    // Operator_Ext_UPML::Create_UPML (slab ranges, operator_ext_upml.cpp:77-209) and
    // BuildExtension (:365-466)
    // UPML slab ranges (Operator_Ext_UPML::Create_UPML) and the slabs' arrays
    std::vector<std::vector<F>> upml_ext;   // [slab*6 + k][3*cnt]
    void UpmlSetup() {
        op->upml.clear();
        upml_ext.clear();
        int* BC = op->BC;
        uint32_t* S = op->pml_size;
        const uint32_t* N = op->N;
        for (int n = 0; n < 3; ++n)
            if ((S[2 * n] * (BC[2 * n] == 3) + S[2 * n + 1] * (BC[2 * n + 1] == 3)) >= N[n]) {
                BC[2 * n] = 0; S[2 * n] = 0; BC[2 * n + 1] = 0; S[2 * n + 1] = 0;
            }
        struct R { uint32_t s[3], e[3]; };
        std::vector<R> ranges;
        uint32_t start[3] = {0, 0, 0}, stop[3] = {N[0] - 1, N[1] - 1, N[2] - 1};
        if (BC[0] == 3) { R r; start[0] = 0; stop[0] = S[0]; memcpy(r.s, start, 12); memcpy(r.e, stop, 12); ranges.push_back(r); }
        if (BC[1] == 3) { R r; start[0] = N[0] - 1 - S[1]; stop[0] = N[0] - 1; memcpy(r.s, start, 12); memcpy(r.e, stop, 12); ranges.push_back(r); }
        start[0] = (S[0] + 1) * (BC[0] == 3);
        stop[0] = N[0] - 1 - (S[0] + 1) * (BC[1] == 3);
        if (BC[2] == 3) { R r; start[1] = 0; stop[1] = S[2]; memcpy(r.s, start, 12); memcpy(r.e, stop, 12); ranges.push_back(r); }
        if (BC[3] == 3) { R r; start[1] = N[1] - 1 - S[3]; stop[1] = N[1] - 1; memcpy(r.s, start, 12); memcpy(r.e, stop, 12); ranges.push_back(r); }
        start[1] = (S[2] + 1) * (BC[2] == 3);
        stop[1] = N[1] - 1 - (S[3] + 1) * (BC[3] == 3);
        if (BC[4] == 3) { R r; start[2] = 0; stop[2] = S[4]; memcpy(r.s, start, 12); memcpy(r.e, stop, 12); ranges.push_back(r); }
        if (BC[5] == 3) { R r; start[2] = N[2] - 1 - S[5]; stop[2] = N[2] - 1; memcpy(r.s, start, 12); memcpy(r.e, stop, 12); ranges.push_back(r); }

        GradingTables(Z0);
        for (const R& r : ranges) {
            wop_operator::Upml u;
            size_t cnt = 1;
            for (int n = 0; n < 3; ++n) { u.start[n] = r.s[n]; u.n[n] = r.e[n] - r.s[n] + 1; cnt *= u.n[n]; }
            op->upml.push_back(std::move(u));
            for (int k = 0; k < 6; ++k) upml_ext.emplace_back(3 * cnt, (F)0);
        }
    }

    // cells of slab si at local plane lx, local row ly
    void UpmlRow(size_t si, uint32_t lx, uint32_t ly) {
        const double dT = op->dT;
        const wop_operator::Upml& u = op->upml[si];
        const size_t cnt = (size_t)u.n[0] * u.n[1] * u.n[2];
        std::vector<F>* ext = &upml_ext[si * 6];
        uint32_t loc[3], pos[3];
        loc[0] = lx;
        loc[1] = ly;
        pos[0] = lx + u.start[0];
        pos[1] = ly + u.start[1];
        double effs[3][4], kv[3], ki[3];
        std::vector<int> cand = op->RowCandidates((int)pos[0], (int)pos[1], P_MATERIAL);
        ZRunReuse run(op, cand, true);
        for (loc[2] = 0; loc[2] < u.n[2]; ++loc[2]) {
            pos[2] = loc[2] + u.start[2];
            const size_t i = op->idx(pos[0], pos[1], pos[2]);
            // the averaged materials repeat where cell z-1's were computed here
            const bool again = run.same(pos[2]) && loc[2] > 0;
            for (int n = 0; n < 3; ++n) {
                const size_t li = (size_t)n * cnt + ((size_t)loc[0] * u.n[1] + loc[1]) * u.n[2] + loc[2];
                double* eff = effs[n];
                if (!again) op->AverageMatQuarterCell(n, pos, eff, cand);
                GradingKappaTab(n, pos, kv, ki);
                const int nP = (n + 1) % 3, nPP = (n + 2) % 3;
                if (((kv[0] + kv[1] + kv[2]) != 0) && (eff[1] < 1e3)) {
                    if ((VV(n, i) + VI(n, i)) != 0) {
                        VV(n, i) = (F)((2 * EPS0 - kv[nP] * dT) / (2 * EPS0 + kv[nP] * dT));
                        VI(n, i) = (F)((2 * EPS0 * dT) / (2 * EPS0 + kv[nP] * dT) * op->EdgeLength(n, pos) / op->EdgeArea(n, pos));
                        ext[0][li] = (F)((2 * EPS0 - kv[nPP] * dT) / (2 * EPS0 + kv[nPP] * dT));
                        ext[2][li] = (F)((2 * EPS0 + kv[n] * dT) / (2 * EPS0 + kv[nPP] * dT) / eff[0]);
                        ext[1][li] = (F)((2 * EPS0 - kv[n] * dT) / (2 * EPS0 + kv[nPP] * dT) / eff[0]);
                    }
                } else {
                    ext[0][li] = VV(n, i);
                    VV(n, i) = 0;
                    ext[1][li] = 0;
                    ext[2][li] = 1;
                }
                if ((ki[0] + ki[1] + ki[2]) != 0) {
                    if ((II(n, i) + IV(n, i)) != 0) {
                        II(n, i) = (F)((2 * EPS0 - ki[nP] * dT) / (2 * EPS0 + ki[nP] * dT));
                        IV(n, i) = (F)((2 * EPS0 * dT) / (2 * EPS0 + ki[nP] * dT) * op->EdgeLength(n, pos, true) / op->EdgeArea(n, pos, true));
                        ext[3][li] = (F)((2 * EPS0 - ki[nPP] * dT) / (2 * EPS0 + ki[nPP] * dT));
                        ext[5][li] = (F)((2 * EPS0 + ki[n] * dT) / (2 * EPS0 + ki[nPP] * dT) / eff[2]);
                        ext[4][li] = (F)((2 * EPS0 - ki[n] * dT) / (2 * EPS0 + ki[nPP] * dT) / eff[2]);
                    }
                } else {
                    ext[3][li] = II(n, i);
                    II(n, i) = 0;
                    ext[4][li] = 0;
                    ext[5][li] = 1;
                }
            }
        }
    }

    void UpmlFinish() {
        for (size_t si = 0; si < op->upml.size(); ++si) {
            wop_operator::Upml& u = op->upml[si];
            for (int k = 0; k < 6; ++k) {
                std::vector<F>& e = upml_ext[si * 6 + k];
                std::vector<double>().swap(u.a[k]);
                std::vector<float>().swap(u.af[k]);
                if constexpr (std::is_same<F, float>::value) u.af[k].swap(e);
                else u.a[k].swap(e);
            }
        }
        upml_ext.clear();
    }

    void Upml() {
        UpmlSetup();
        for (size_t si = 0; si < op->upml.size(); ++si) {
            const wop_operator::Upml& u = op->upml[si];
            op->parallel_x(u.n[0], [&](uint32_t lx) {
                for (uint32_t ly = 0; ly < u.n[1]; ++ly) UpmlRow(si, lx, ly);
            });
        }
        UpmlFinish();
    }

    void UpmlPlane(uint32_t x) {
        // the slabs cover disjoint cells: one parallel sweep over the rows of
        // the plane handles every slab that the row crosses
        std::vector<size_t> here;
        for (size_t si = 0; si < op->upml.size(); ++si) {
            const wop_operator::Upml& u = op->upml[si];
            if (x >= u.start[0] && x < u.start[0] + u.n[0]) here.push_back(si);
        }
        if (here.empty()) return;
        op->parallel_x(op->N[1], [&](uint32_t y) {
            for (size_t si : here) {
                const wop_operator::Upml& u = op->upml[si];
                if (y >= u.start[1] && y < u.start[1] + u.n[1]) UpmlRow(si, x - u.start[0], y - u.start[1]);
            }
        });
    }

    // Operator::CalcNumericPhaseVelocity (FDTD/operator.cpp:2087-2151)
    double NumericPhaseVelocity(const uint32_t* start, const uint32_t* stop, const double* dir, float freq) const {
        const double dT = op->dT;
        const double c0 = C0;   // background eps_r = mue_r = 1
        double disc[3];
        for (int n = 0; n < 3; ++n)
            disc[n] = std::fabs(op->DiscLine(n, start[n], false) - op->DiscLine(n, stop[n], false)) * op->gridDelta /
                      std::fabs((double)(stop[n] - start[n]));
        for (int n = 0; n < 3; ++n) {
            const int nP = (n + 1) % 3, nPP = (n + 2) % 3;
            if (std::fabs(dir[n]) == 1 && dir[nP] == 0 && dir[nPP] == 0) {
                // the reference takes the x spacing for every Cartesian direction
                double kx = asin(disc[0] / c0 / dT * sin(2 * PI_ * freq * dT / 2)) * 2 / disc[0];
                return 2 * PI_ * freq / kx;
            }
        }
        const double k0 = 2 * PI_ * freq / c0;
        double k = k0;
        const double RHS = pow(sin(2 * PI_ * freq * dT / 2) / c0 / dT, 2);
        double old_phv = 0, phv = c0;
        int it = 0;
        while (std::fabs(old_phv - phv) > 1e-6) {
            ++it;
            old_phv = phv;
            double fk = 0, fdk = 0;
            for (int n = 0; n < 3; ++n) {
                fk += pow(sin(dir[n] * k * disc[n] / 2) / disc[n], 2);
                fdk += dir[n] * sin(dir[n] * k * disc[n] / 2) * cos(dir[n] * k * disc[n] / 2) / disc[n];
            }
            fk -= RHS;
            k -= fk / fdk;
            if (k < k0) k = k0;
            phv = 2 * PI_ * freq / k;
            if (it > 99) break;
        }
        return phv;
    }

    // Operator_Ext_TFSF::BuildExtension (operator_ext_tfsf.cpp:82-410).  The
    // delay, fraction and amplitude tables are built in the reference's
    // statement order (F rounding at its storage points) and flattened in the
    // order Engine_Ext_TFSF applies them (engine_ext_tfsf.cpp:36-215): axis n,
    // lower then upper plane, nP-major then nPP, component nP then nPP.
    void Tfsf() {
        op->tfsf_v.clear();
        op->tfsf_i.clear();
        op->tfsf_max_delay = 0;
        const auto& pw = op->pw;
        if (!pw.set) return;
        const double dT = op->dT;
        const uint32_t* N = op->N;
        double k[3] = {pw.k[0], pw.k[1], pw.k[2]};
        const double norm = sqrt(k[0] * k[0] + k[1] * k[1] + k[2] * k[2]);
        if (norm == 0) return;
        k[0] /= norm; k[1] /= norm; k[2] /= norm;
        // SnapBox2Mesh (FDTD/operator.cpp:289-314), snap method 0, main mesh
        uint32_t S[3], T[3];
        int dims = 0;
        for (int n = 0; n < 3; ++n) {
            const double lo = fmin(pw.a[n], pw.b[n]), hi = fmax(pw.a[n], pw.b[n]);
            const double mn = op->DiscLine(n, 0, false), mx = op->DiscLine(n, N[n] - 1, false);
            if ((lo < mn && hi < mn) || (lo > mx && hi > mx)) return;
        }
        for (int n = 0; n < 3; ++n) {
            bool in;
            S[n] = op->SnapToMeshLine(n, fmin(pw.a[n], pw.b[n]), in, false);
            T[n] = op->SnapToMeshLine(n, fmax(pw.a[n], pw.b[n]), in, false);
            if (T[n] > S[n]) ++dims;
        }
        if (dims != 3) return;
        double phv = C0;
        if (pw.f > 0) phv = NumericPhaseVelocity(S, T, k, (float)pw.f);
        if (phv < 0 || phv > C0 || std::isnan(phv)) phv = C0;

        double origin[3];
        uint32_t nl[3];
        bool active[3][2];
        double E[3] = {pw.E[0], pw.E[1], pw.E[2]}, H[3];
        for (int n = 0; n < 3; ++n) {
            nl[n] = T[n] - S[n] + 1;
            active[n][0] = S[n] != 0;
            active[n][1] = T[n] != N[n] - 1;
            const uint32_t ui = k[n] >= 0 ? S[n] - 1 : T[n] + 1;   // wraps at the domain edge, as in the reference
            origin[n] = op->DiscLine(n, ui, false);
        }
        const double dotEk = E[0] * k[0] + E[1] * k[1] + E[2] * k[2];
        const double angle = acos(dotEk / (E[0] * E[0] + E[1] * E[1] + E[2] * E[2])) / PI_ * 180;
        if (angle == 0) return;
        if (angle != 90)
            for (int n = 0; n < 3; ++n) E[n] -= k[n] * dotEk;
        for (int n = 0; n < 3; ++n) {
            const int nP = (n + 1) % 3, nPP = (n + 2) % 3;
            H[n] = k[nP] * E[nPP] - k[nPP] * E[nP];
            H[n] /= Z0 * sqrt(1.0 / 1.0);
        }
        const double unit = op->gridDelta;
        uint32_t maxDelay = 0;
        auto delay_of = [&](const uint32_t* pos, int comp, bool dual, double extra) {
            double c[3];
            op->YeeCoords(comp, pos, c, dual);
            const double dist = std::fabs((c[0] - origin[0]) * k[0]) + std::fabs((c[1] - origin[1]) * k[1]) +
                                std::fabs((c[2] - origin[2]) * k[2]);
            const double d = dist * unit / phv / dT + extra;
            maxDelay = std::max((uint32_t)d, maxDelay);
            return d;
        };
        struct Tab { std::vector<uint32_t> delay; std::vector<F> delta, amp; };
        for (int n = 0; n < 3; ++n) {
            const int nP = (n + 1) % 3, nPP = (n + 2) % 3;
            if (!active[n][0] && !active[n][1]) continue;
            const size_t numP = (size_t)nl[nP] * nl[nPP];
            Tab Vt[2][2], Ct[2][2];
            for (int l = 0; l < 2; ++l)
                for (int c = 0; c < 2; ++c)
                    if (active[n][l]) {
                        Vt[l][c].delay.assign(numP, 0); Vt[l][c].delta.assign(numP, 0); Vt[l][c].amp.assign(numP, 0);
                        Ct[l][c].delay.assign(numP, 0); Ct[l][c].delta.assign(numP, 0); Ct[l][c].amp.assign(numP, 0);
                    }
            auto put = [](Tab& t, size_t u, double d) {
                t.delay[u] = (uint32_t)floor(d);
                t.delta[u] = (F)(d - floor(d));
            };
            uint32_t pos[3];
            pos[n] = 0;
            pos[nP] = S[nP];
            size_t u = 0;
            for (uint32_t i = 0; i < nl[nP]; ++i) {
                pos[nPP] = S[nPP];
                for (uint32_t j = 0; j < nl[nPP]; ++j) {
                    // current injection tables
                    pos[n] = S[n];
                    if (active[n][0]) {
                        put(Ct[0][1], u, delay_of(pos, nP, false, 0.0));
                        Ct[0][1].amp[u] = (F)(E[nP] * op->EdgeLength(nP, pos));
                        put(Ct[0][0], u, delay_of(pos, nPP, false, 0.0));
                        Ct[0][0].amp[u] = (F)(E[nPP] * op->EdgeLength(nPP, pos));
                        --pos[n];
                        const size_t q = op->idx(pos[0], pos[1], pos[2]);
                        Ct[0][0].amp[u] = Ct[0][0].amp[u] * IV(nP, q);
                        Ct[0][1].amp[u] = Ct[0][1].amp[u] * IV(nPP, q);
                    }
                    if (active[n][1]) {
                        pos[n] = T[n];
                        put(Ct[1][1], u, delay_of(pos, nP, false, 0.0));
                        Ct[1][1].amp[u] = (F)(E[nP] * op->EdgeLength(nP, pos));
                        put(Ct[1][0], u, delay_of(pos, nPP, false, 0.0));
                        Ct[1][0].amp[u] = (F)(E[nPP] * op->EdgeLength(nPP, pos));
                        const size_t q = op->idx(pos[0], pos[1], pos[2]);
                        Ct[1][0].amp[u] = Ct[1][0].amp[u] * IV(nP, q);
                        Ct[1][1].amp[u] = Ct[1][1].amp[u] * IV(nPP, q);
                    }
                    if (active[n][0]) Ct[0][0].amp[u] = Ct[0][0].amp[u] * (F)-1;
                    if (active[n][1]) Ct[1][1].amp[u] = Ct[1][1].amp[u] * (F)-1;
                    if (pos[nP] == T[nP]) {
                        if (active[n][0]) Ct[0][1].amp[u] = 0;
                        if (active[n][1]) Ct[1][1].amp[u] = 0;
                    }
                    if (pos[nPP] == T[nPP]) {
                        if (active[n][0]) Ct[0][0].amp[u] = 0;
                        if (active[n][1]) Ct[1][0].amp[u] = 0;
                    }
                    // voltage injection tables
                    pos[n] = S[n] - 1;
                    if (active[n][0]) {
                        put(Vt[0][1], u, delay_of(pos, nP, true, 1.0));
                        Vt[0][1].amp[u] = (F)(H[nP] * op->EdgeLength(nP, pos, true));
                        put(Vt[0][0], u, delay_of(pos, nPP, true, 1.0));
                        Vt[0][0].amp[u] = (F)(H[nPP] * op->EdgeLength(nPP, pos, true));
                        ++pos[n];
                        const size_t q = op->idx(pos[0], pos[1], pos[2]);
                        Vt[0][0].amp[u] = Vt[0][0].amp[u] * VI(nP, q);
                        Vt[0][1].amp[u] = Vt[0][1].amp[u] * VI(nPP, q);
                    }
                    pos[n] = T[n];
                    if (active[n][1]) {
                        put(Vt[1][1], u, delay_of(pos, nP, true, 1.0));
                        Vt[1][1].amp[u] = (F)(H[nP] * op->EdgeLength(nP, pos, true));
                        put(Vt[1][0], u, delay_of(pos, nPP, true, 1.0));
                        Vt[1][0].amp[u] = (F)(H[nPP] * op->EdgeLength(nPP, pos, true));
                        const size_t q = op->idx(pos[0], pos[1], pos[2]);
                        Vt[1][0].amp[u] = Vt[1][0].amp[u] * VI(nP, q);
                        Vt[1][1].amp[u] = Vt[1][1].amp[u] * VI(nPP, q);
                    }
                    if (active[n][1]) Vt[1][0].amp[u] = Vt[1][0].amp[u] * (F)-1;
                    if (active[n][0]) Vt[0][1].amp[u] = Vt[0][1].amp[u] * (F)-1;
                    if (pos[nP] == T[nP]) {
                        if (active[n][0]) Vt[0][0].amp[u] = 0;
                        if (active[n][1]) Vt[1][0].amp[u] = 0;
                    }
                    if (pos[nPP] == T[nPP]) {
                        if (active[n][0]) Vt[0][1].amp[u] = 0;
                        if (active[n][1]) Vt[1][1].amp[u] = 0;
                    }
                    ++pos[nPP];
                    ++u;
                }
                ++pos[nP];
            }
            // flatten in the engine's application order
            for (int l = 0; l < 2; ++l) {
                if (!active[n][l]) continue;
                for (int which = 0; which < 2; ++which) {
                    Tab (&tb)[2][2] = which == 0 ? Vt : Ct;
                    auto& out = which == 0 ? op->tfsf_v : op->tfsf_i;
                    uint32_t p[3];
                    size_t uu = 0;
                    for (uint32_t i = 0; i < nl[nP]; ++i)
                        for (uint32_t j = 0; j < nl[nPP]; ++j, ++uu) {
                            p[nP] = S[nP] + i;
                            p[nPP] = S[nPP] + j;
                            p[n] = l == 0 ? (which == 0 ? S[n] : S[n] - 1) : T[n];
                            for (int c = 0; c < 2; ++c) {
                                wop_operator::TfsfEntry e;
                                for (int a = 0; a < 3; ++a) e.pos[a] = p[a];
                                e.comp = c == 0 ? nP : nPP;
                                e.delay = tb[l][c].delay[uu];
                                e.delta = (double)tb[l][c].delta[uu];
                                e.amp = (double)tb[l][c].amp[uu];
                                out.push_back(e);
                            }
                        }
                }
            }
        }
        op->tfsf_max_delay = maxDelay + 1;
    }

    // Operator_Ext_LorentzMaterial::BuildExtension (operator_ext_lorentzmaterial.cpp:117-438),
    // order 1, electric Drude/Lorentz poles
    void Lorentz() {
        op->ade.clear();
        bool any = false;
        for (auto& p : op->props) if (p.lorentz) any = true;
        if (!any) return;
        const uint32_t* N = op->N;
        const double dT = op->dT;
        wop_operator::Ade a;
        bool volt_on = false, volt_lor_on = false;
        std::vector<double> v_int[3], v_ext[3], v_lor[3];
        uint32_t pos[3];
        double c[3];
        for (pos[0] = 0; pos[0] < N[0]; ++pos[0])
            for (pos[1] = 0; pos[1] < N[1]; ++pos[1]) {
                std::vector<int> cand = op->RowCandidates((int)pos[0], (int)pos[1], P_MATERIAL | P_METAL);
                for (pos[2] = 0; pos[2] < N[2]; ++pos[2]) {
                    const size_t index = op->idx(pos[0], pos[1], pos[2]);
                    bool on = false;
                    double L_D[3], R_D[3], C_L[3];
                    for (int n = 0; n < 3; ++n) {
                        L_D[n] = 0; R_D[n] = 0; C_L[n] = 0;
                        if (!op->YeeCoords(n, pos, c, false)) continue;
                        if (VI(n, index) == 0) continue;
                        int p = op->PropertyAt(c, P_MATERIAL | P_METAL, &cand);
                        if (p < 0 || !op->props[p].lorentz) continue;
                        const Property& P = op->props[p];
                        double w_plasma = P.eps_plasma_f * 2 * PI_;
                        if ((w_plasma > 0) && (EC[0][n][index] > 0)) {
                            on = true;
                            volt_on = true;
                            L_D[n] = 1 / (w_plasma * w_plasma * (double)EC[0][n][index]);
                        }
                        double t_relax = P.eps_relax_t;
                        if ((t_relax > 0) && volt_on) R_D[n] = L_D[n] / t_relax;
                        double w_lor = P.eps_lor_f * 2 * PI_;
                        if ((w_lor > 0) && (L_D[n] > 0)) {
                            volt_lor_on = true;
                            C_L[n] = 1 / (w_lor * w_lor * L_D[n]);
                        }
                    }
                    if (!on) continue;
                    for (int n = 0; n < 3; ++n) {
                        a.pos[n].push_back(pos[n]);
                        const double vi = (double)VI(n, index);
                        if (L_D[n] > 0) {
                            v_int[n].push_back((double)(F)((2.0 * L_D[n] - dT * R_D[n]) / (2.0 * L_D[n] + dT * R_D[n])));
                            v_ext[n].push_back((double)(F)(dT / (L_D[n] + dT * R_D[n] / 2.0) * vi));
                        } else if ((R_D[n] > 0) && (C_L[n] > 0)) {
                            v_int[n].push_back((double)(F)((2.0 * dT - R_D[n] * C_L[n]) / (C_L[n] * R_D[n])));
                            v_ext[n].push_back((double)(F)(2.0 / R_D[n] * vi));
                        } else {
                            v_int[n].push_back(1);
                            v_ext[n].push_back(0);
                        }
                        if (C_L[n] > 0) v_lor[n].push_back((double)(F)(dT / C_L[n] / vi));
                        else v_lor[n].push_back(0);
                    }
                }
            }
        const size_t cnt = a.pos[0].size();
        if (!cnt) return;
        for (int k = 0; k < 6; ++k) a.lists[k].assign(3 * cnt, 0.0);
        for (int n = 0; n < 3; ++n)
            for (size_t i = 0; i < cnt; ++i) {
                a.lists[0][n * cnt + i] = v_int[n][i];
                a.lists[1][n * cnt + i] = v_ext[n][i];
                a.lists[2][n * cnt + i] = v_lor[n][i];
                a.lists[3][n * cnt + i] = 1;
                a.lists[4][n * cnt + i] = 0;
            }
        a.flags = (volt_on ? 1u : 0u) | (volt_lor_on ? 4u : 0u);
        op->ade.push_back(std::move(a));
    }

    // Excitation::CalcGaussianPulsExcitation etc. (FDTD/excitation.cpp:151-277)
    // Operator_Ext_LumpedRLC::BuildExtension (operator_ext_lumpedRLC.cpp:110-452):
    // series RLC and parallel RLC with an inductance.  Each element's cells get
    // the per-cell C (parallel: C or the cell's own capacitance; series: the
    // cell's capacitance raised to keep its impedance 20x above the element's
    // at fMax) and G, their main coefficients are recomputed, and the ADE
    // coefficient lists are appended in the reference's loop order.  Runs
    // after UPML and Lorentz (the extension is created last, openems.cpp:1096).
    void LumpedRLC() {
        auto& out = op->rlc;
        for (int k = 0; k < 3; ++k) out.pos[k].clear();
        out.dir.clear();
        for (auto& v : out.c) v.clear();
        const double dT = op->dT;
        const double fMax = op->f0 + op->fc;           // GetCenterFreq() + GetCutOffFreq()
        const double ZF = 20.0;                        // LUMPED_RLC_Z_FACT (operator_ext_lumpedRLC.h:29)
        for (const Primitive& pr : op->prims) {
            const Property& P = op->props[pr.prop];
            if (!(P.kind & P_LUMPED) || pr.shape != 0) continue;
            // IsLElumpedRLC (operator_ext_lumpedRLC.cpp:500-512)
            const bool parallel = P.le_type == 0;
            if (!((parallel && P.L > 0.0) || P.le_type == 1)) continue;
            double C = P.C, R = P.R, L = P.L;
            if (C < 0) C = NAN;
            if (R < 0) R = NAN;
            if (L < 0) L = NAN;
            const int dir = P.ny;
            if (dir < 0 || dir > 2) continue;
            const int d1 = (dir + 1) % 3, d2 = (dir + 2) % 3;
            uint32_t st[3], sp[3];
            bool in;
            for (int n = 0; n < 3; ++n) {
                st[n] = op->SnapToMeshLine(n, pr.a[n], in, false);
                sp[n] = op->SnapToMeshLine(n, pr.b[n], in, false);
            }
            if (st[dir] == sp[dir]) continue;
            const unsigned n0 = sp[dir] - st[dir], n1 = sp[d1] - st[d1] + 1, n2 = sp[d2] - st[d2] + 1;
            const unsigned npar = n1 * n2;
            const double dL = L * npar / n0, dR = R * npar / n0;
            double dG = (1 / R) * n0 / npar, dC = C * n0 / npar;
            const double ib0 = 2.0 * dT * dC / (4.0 * dL * dC + 2.0 * dT * dR * dC + dT * dT),
                         b1 = (dT * dT - 4.0 * dL * dC) / (dT * dC),
                         b2 = (4.0 * dL * dC - 2.0 * dT * dR * dC + dT * dT) / (2.0 * dT * dC);
            if (parallel && R == 0.0) dG = 0.0;
            uint32_t pos[3];
            for (pos[dir] = st[dir]; pos[dir] < sp[dir]; ++pos[dir])
                for (pos[d1] = st[d1]; pos[d1] <= sp[d1]; ++pos[d1])
                    for (pos[d2] = st[d2]; pos[d2] <= sp[d2]; ++pos[d2]) {
                        const size_t i = op->idx(pos[0], pos[1], pos[2]);
                        double c[9];
                        if (parallel) {
                            if (dC > 0) EC[0][dir][i] = (F)dC;
                            else dC = (double)EC[0][dir][i];      // the first cell's own capacitance, kept for the rest
                            c[1] = (dT / dC) / (1.0 + dT * dG / (2.0 * dC));
                            if (R >= 0) EC[1][dir][i] = (F)dG;
                            c[0] = L > 0 ? dT / dL : 0.0;
                            if (dC == 0) {
                                double Cd = (double)EC[0][dir][i];
                                const double Zmin = std::max(dR, 2 * PI_ * fMax * dL);
                                const double Zcd_min = 1.0 / (2.0 * PI_ * fMax * Cd);
                                if (Zcd_min < ZF * Zmin) {
                                    Cd = 1.0 / (2 * PI_ * fMax * Zmin * ZF);
                                    EC[0][dir][i] = (F)Cd;
                                }
                            }
                            c[2] = 0.0; c[3] = 0.0; c[4] = 0.0; c[5] = 1.0; c[6] = 0.0; c[7] = 0.0; c[8] = 0.0;
                        } else {
                            EC[1][dir][i] = (F)0.0;
                            F Cd = EC[0][dir][i];
                            const double Zmin = sqrt(pow(dR, 2) + pow(2 * PI_ * fMax * dL - 1.0 / (dC * 2 * PI_ * fMax), 2));
                            const double Zcd_min = 1.0 / (2.0 * PI_ * fMax * Cd);
                            if (Zcd_min < ZF * Zmin) {
                                Cd = (F)(1.0 / (2 * PI_ * fMax * Zmin * ZF));
                                EC[0][dir][i] = Cd;
                            }
                            c[0] = 0.0; c[1] = 0.0;
                            c[2] = 0.5 * dT * ib0 / Cd;
                            c[3] = 0.5 * dT * (b1 * ib0 - 1.0) / Cd;
                            c[4] = 0.5 * dT * b2 * ib0 / Cd;
                            c[5] = 1.0 / (1.0 + 0.5 * dT * ib0 / Cd);
                            c[6] = ib0; c[7] = b1; c[8] = b2;
                        }
                        ECOperatorPos(dir, i);
                        for (int k = 0; k < 9; ++k) out.c[k].push_back((double)(F)c[k]);
                        out.dir.push_back((uint32_t)dir);
                        for (int k = 0; k < 3; ++k) out.pos[k].push_back(pos[k]);
                    }
            if (P.caps) {
                for (pos[d1] = st[d1]; pos[d1] <= sp[d1]; ++pos[d1])
                    for (pos[d2] = st[d2]; pos[d2] <= sp[d2]; ++pos[d2])
                        for (int side = 0; side < 2; ++side) {
                            pos[dir] = side ? sp[dir] : st[dir];
                            const size_t i = op->idx(pos[0], pos[1], pos[2]);
                            if (pos[d1] < sp[d1]) { VV(d1, i) = 0; VI(d1, i) = 0; }
                            if (pos[d2] < sp[d2]) { VV(d2, i) = 0; VI(d2, i) = 0; }
                        }
            }
        }
    }

    // Operator_Ext_Mur_ABC::SetDirection/BuildExtension (operator_ext_mur_abc.cpp:88-180)
    // and the excitation start rule of Engine_Ext_Mur_ABC (engine_ext_mur_abc.cpp:42-58).
    // Needs the excitation lists and the signal length (call after Signal).
    void Mur() {
        op->mur.clear();
        const uint32_t* N = op->N;
        const double dT = op->dT;
        for (int face = 0; face < 6; ++face) {
            if (op->BC[face] != WOP_BC_MUR) continue;
            const int ny = face / 2, nyP = (ny + 1) % 3, nyPP = (ny + 2) % 3;
            const bool top = face % 2;
            const uint32_t line = top ? N[ny] - 1 : 0;
            wop_operator::Mur m;
            m.face = face;
            m.na = N[nyP];
            m.nb = N[nyPP];
            m.kP.assign((size_t)m.na * m.nb, 0.0);
            m.kPP.assign((size_t)m.na * m.nb, 0.0);
            uint32_t pos[3] = {0, 0, 0};
            pos[ny] = line;
            const double delta = std::fabs(op->EdgeLength(ny, pos));
            double coord[3];
            for (int n = 0; n < 3; ++n) coord[n] = op->DiscLine(n, pos[n], false);
            if (line == 0) coord[ny] = op->DiscLine(ny, pos[ny], false) + delta / 2 / op->gridDelta;
            else coord[ny] = op->DiscLine(ny, pos[ny], false) - delta / 2 / op->gridDelta;
            const double vph = op->mur_vph[face];
            for (uint32_t a = 0; a < m.na; ++a) {
                coord[nyP] = op->DiscLine(nyP, a, false);
                for (uint32_t b = 0; b < m.nb; ++b) {
                    coord[nyPP] = op->DiscLine(nyPP, b, false);
                    const int p = op->PropertyAt(coord, P_MATERIAL);
                    const size_t i = (size_t)a * m.nb + b;
                    double c0t;
                    if (p >= 0) {
                        const Property& P = op->props[p];
                        c0t = vph > 0.0 ? vph * dT : C0 * dT / sqrt(P.eps * P.mue);
                        m.kP[i] = (double)(F)((c0t - delta) / (c0t + delta));
                        c0t = vph > 0.0 ? vph * dT : C0 * dT / sqrt(P.eps * P.mue);
                        m.kPP[i] = (double)(F)((c0t - delta) / (c0t + delta));
                    } else {
                        // background material eps_r = mue_r = 1
                        c0t = vph > 0.0 ? vph * dT : C0 / sqrt(1.0 * 1.0) * dT;
                        m.kP[i] = (double)(F)((c0t - delta) / (c0t + delta));
                        m.kPP[i] = m.kP[i];
                    }
                }
            }
            int max_delay = -1;
            for (const auto& e : op->exc_v)
                if (((int)e.dir == nyP || (int)e.dir == nyPP) && e.pos[ny] == line && (int)e.delay > max_delay)
                    max_delay = (int)e.delay;
            m.start_ts = max_delay >= 0 ? (uint32_t)max_delay + op->exc_length + 10 : 0;
            op->mur.push_back(std::move(m));
        }
    }

    void Signal(uint32_t nr_ts) {
        const double dT = op->dT;
        std::vector<F> sv, si;
        uint32_t L = 0;
        double fmax = 0;
        switch (op->exc_sig_type) {
        case WOP_EXC_GAUSS: {
            const double f0 = op->f0, fc = op->fc;
            L = (uint32_t)ceil(2.0 * 9.0 / (2.0 * PI_ * fc) / dT);
            if (L > nr_ts) L = nr_ts;
            sv.resize(L); si.resize(L);
            for (uint32_t n = 0; n < L; ++n) {
                double t = n * dT;
                sv[n] = (F)(cos(2.0 * PI_ * f0 * (t - 9.0 / (2.0 * PI_ * fc))) * exp(-1 * pow(2.0 * PI_ * fc * t / 3.0 - 3, 2)));
                t += 0.5 * dT;
                si[n] = (F)(cos(2.0 * PI_ * f0 * (t - 9.0 / (2.0 * PI_ * fc))) * exp(-1 * pow(2.0 * PI_ * fc * t / 3.0 - 3, 2)));
            }
            fmax = f0 + fc;
            break;
        }
        case WOP_EXC_SINE: {
            const double f0 = op->f0;
            L = (uint32_t)round(2.0 / f0 / dT);
            sv.assign(L, (F)0); si.assign(L, (F)0);
            for (uint32_t n = 1; n < L; ++n) {
                double t = n * dT;
                sv[n] = (F)sin(2.0 * PI_ * f0 * t);
                t += 0.5 * dT;
                si[n] = (F)sin(2.0 * PI_ * f0 * t);
            }
            fmax = f0;
            break;
        }
        case WOP_EXC_DIRAC:
            L = 2; sv = {(F)0, (F)1}; si = {(F)0, (F)1}; fmax = -1;
            break;
        default:
            L = 2; sv = {(F)1, (F)1}; si = {(F)1, (F)1}; fmax = -1;
            break;
        }
        op->sig_v.assign(sv.begin(), sv.end());
        op->sig_i.assign(si.begin(), si.end());
        op->exc_length = L;
        // CalcNyquistNum, tools/useful.cpp:30-36
        if (fmax < 0) op->nyquist = 1;
        else if (fmax == 0) op->nyquist = 0xffffffffu;
        else op->nyquist = (uint32_t)floor(1 / fmax / 2 / dT);
        // GetMaxExcitationTimestep (FDTD/excitation.cpp:136-149)
        F maxAmp = 0;
        uint32_t maxStep = 0;
        for (uint32_t n = 0; n < L; ++n)
            if (std::fabs(sv[n]) > maxAmp) { maxAmp = std::fabs(sv[n]); maxStep = n; }
        op->max_step = maxStep;
    }
};

}  // namespace

template <> Store<float>& wop_operator::store<float>() { return sf; }
template <> Store<double>& wop_operator::store<double>() { return sd; }

template <typename F>
int wop_operator::build(uint32_t nr_ts) {
    sf.clear();
    sd.clear();
    FillEdgeTables();
    InitMaterialStorage();
    ring = 0;
    streamed = false;
    stream_table.clear();
    Builder<F> b(this);
    b.CalcEC();
    // CalcECOperator timestep handling (FDTD/operator.cpp:981-1007)
    if (forced_dT > 0) dT = forced_dT;
    else dT = b.CalcTimestep();
    dT *= ts_factor;
    if (exc_sig_type == WOP_EXC_SINE && f0 > 0) {
        double period = 1 / f0;
        unsigned TS = (unsigned)ceil(period / dT);
        dT = period / TS;
    }
    const size_t n = nd();
    for (int k = 0; k < 4; ++k)
        for (int d = 0; d < 3; ++d) b.co[k][d].assign(n, (F)0);
    parallel_x(N[0], [&](uint32_t x) {
        for (int d = 0; d < 3; ++d)
            for (size_t i = (size_t)x * N[1] * N[2]; i < (size_t)(x + 1) * N[1] * N[2]; ++i) b.ECOperatorPos(d, i);
    });
    b.ApplyElectricBC();
    b.CalcPEC();
    b.LumpedElements();
    b.ApplyMagneticBC();
    // extensions in creation order: excitation, (TFSF), UPML, Lorentz (openems.cpp:1041-1092)
    b.Excitation();
    b.Tfsf();
    b.Upml();
    b.Lorentz();
    b.LumpedRLC();
    b.Signal(nr_ts);
    b.Mur();
    built = true;
    return 0;
}

template <typename F>
static void put3(const std::vector<F> (&src)[3], int dtype, void* dst) {
    if (!dst) return;
    const size_t n = src[0].size();
    for (int d = 0; d < 3; ++d) {
        if (dtype == 1) {
            double* o = (double*)dst + d * n;
            for (size_t i = 0; i < n; ++i) o[i] = (double)src[d][i];
        } else {
            float* o = (float*)dst + d * n;
            for (size_t i = 0; i < n; ++i) o[i] = (float)src[d][i];
        }
    }
}

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int wop_create(wop_operator** out) {
    if (!out) return fail(-1, "null argument");
    *out = new wop_operator();
    return 0;
}

int wop_destroy(wop_operator* op) {
    delete op;
    return 0;
}

int wop_last_error(char* buf, size_t len) {
    if (!buf || !len) return -1;
    snprintf(buf, len, "%s", g_err.c_str());
    return 0;
}

int wop_set_mesh(wop_operator* op, int axis, const double* lines, uint32_t n) {
    if (!op || !lines || axis < 0 || axis > 2) return fail(-1, "bad argument");
    if (n < 2) return fail(-1, "need at least 2 mesh lines");
    for (uint32_t i = 1; i < n; ++i)
        if (!(lines[i] > lines[i - 1])) return fail(-1, "mesh lines must increase");
    op->lines[axis].assign(lines, lines + n);
    op->N[axis] = n;
    op->built = false;
    op->xcand.clear();
    for (int dl = 0; dl < 2; ++dl) op->edge_len[dl][axis].clear();
    return 0;
}

int wop_set_grid_delta(wop_operator* op, double unit) {
    if (!op || !(unit > 0)) return fail(-1, "bad argument");
    op->gridDelta = unit;
    for (int dl = 0; dl < 2; ++dl)
        for (int n = 0; n < 3; ++n) op->edge_len[dl][n].clear();
    return 0;
}

int wop_set_boundary(wop_operator* op, const int bc[6], const uint32_t pml_size[6]) {
    if (!op || !bc) return fail(-1, "bad argument");
    for (int k = 0; k < 6; ++k) {
        if (bc[k] < -1 || bc[k] > 3) return fail(-1, "bad boundary code %d", bc[k]);
        op->BC[k] = bc[k];
        op->pml_size[k] = (bc[k] == WOP_BC_PML && pml_size) ? pml_size[k] : 0;
    }
    return 0;
}

int wop_set_mur_phase_velocity(wop_operator* op, int face, double v) {
    if (!op || face < 0 || face > 5 || v < 0) return fail(-1, "bad argument");
    op->mur_vph[face] = v;
    return 0;
}

int wop_get_num_mur(wop_operator* op, uint32_t* n) {
    if (!op || !n || !op->built) return fail(-1, "operator not built");
    *n = (uint32_t)op->mur.size();
    return 0;
}

int wop_get_mur(wop_operator* op, uint32_t k, int* face, int dtype, void* k_nyP, void* k_nyPP, uint32_t* start_ts) {
    if (!op || !op->built || k >= op->mur.size() || !face || !start_ts) return fail(-1, "bad argument");
    const wop_operator::Mur& m = op->mur[k];
    *face = m.face;
    *start_ts = m.start_ts;
    const std::vector<double>* src[2] = {&m.kP, &m.kPP};
    void* dst[2] = {k_nyP, k_nyPP};
    for (int j = 0; j < 2; ++j) {
        if (!dst[j]) continue;
        for (size_t i = 0; i < src[j]->size(); ++i) {
            if (dtype == 1) ((double*)dst[j])[i] = (*src[j])[i];
            else ((float*)dst[j])[i] = (float)(*src[j])[i];
        }
    }
    return 0;
}

int wop_set_timestep_factor(wop_operator* op, double f) {
    if (!op || f <= 0 || f > 1) return fail(-1, "invalid timestep factor");
    op->ts_factor = f;
    return 0;
}

int wop_set_timestep(wop_operator* op, double dT) {
    if (!op || dT < 0) return fail(-1, "bad argument");
    op->forced_dT = dT;
    return 0;
}

int wop_set_threads(wop_operator* op, int n) {
    if (!op) return fail(-1, "bad argument");
    op->nthreads = n;
    return 0;
}

int wop_set_verbose(wop_operator* op, int level) {
    if (!op) return fail(-1, "bad argument");
    op->verbose = level;
    return 0;
}

int wop_set_row_reuse(wop_operator* op, int on) {
    if (!op) return fail(-1, "bad argument");
    op->row_reuse = on != 0;
    return 0;
}

static int add_prop(wop_operator* op, const Property& p, int* id) {
    op->props.push_back(p);
    if (id) *id = (int)op->props.size() - 1;
    return 0;
}

int wop_add_material(wop_operator* op, double eps, double kappa, double mue, double sigma, int* id) {
    if (!op) return fail(-1, "bad argument");
    Property p;
    p.kind = P_MATERIAL; p.eps = eps; p.kappa = kappa; p.mue = mue; p.sigma = sigma;
    return add_prop(op, p, id);
}

int wop_set_density(wop_operator* op, int prop, double density) {
    if (!op) return fail(-1, "bad argument");
    if (prop < 0 || prop >= (int)op->props.size() || !(op->props[prop].kind & P_MATERIAL))
        return fail(-1, "property %d is not a material", prop);
    if (!(density >= 0)) return fail(-1, "density must be >= 0");
    op->props[prop].density = density;
    return 0;
}

int wop_get_cell_materials(wop_operator* op, const uint32_t start[3], const uint32_t stop[3], float* kappa, float* density, float* volume) {
    if (!op || !start || !stop) return fail(-1, "bad argument");
    if (op->lines[0].empty() || op->lines[1].empty() || op->lines[2].empty()) return fail(-1, "mesh not set");
    for (int n = 0; n < 3; ++n)
        if (start[n] > stop[n] || stop[n] >= op->N[n]) return fail(-1, "box out of the mesh on axis %d", n);
    const uint32_t nx = stop[0] - start[0] + 1, ny = stop[1] - start[1] + 1, nz = stop[2] - start[2] + 1;
    op->parallel_x(nx, [&](uint32_t i) {
        uint32_t pos[3];
        pos[0] = start[0] + i;
        for (uint32_t j = 0; j < ny; ++j) {
            pos[1] = start[1] + j;
            for (uint32_t k = 0; k < nz; ++k) {
                pos[2] = start[2] + k;
                const size_t q = ((size_t)i * ny + j) * nz + k;
                // the material at the cell centre (Operator::GetCellCenterMaterialAvgCoord,
                // FDTD/operator.cpp:1238-1249, then GetPropertyByCoordPriority over the
                // material primitives, Common/processfields_sar.cpp:200-230)
                double c[3];
                for (int n = 0; n < 3; ++n) c[n] = op->DiscLine(n, pos[n], true);
                const int pr = op->PropertyAt(c, P_MATERIAL);
                if (kappa) kappa[q] = pr >= 0 ? (float)op->props[pr].kappa : 0.0f;
                if (density) density[q] = pr >= 0 ? (float)op->props[pr].density : 0.0f;
                // Operator::GetCellVolume (FDTD/operator.cpp:209-215)
                if (volume) {
                    double v = 1;
                    for (int n = 0; n < 3; ++n) v *= op->EdgeLength(n, pos, false);
                    volume[q] = (float)v;
                }
            }
        }
    });
    return 0;
}

int wop_add_metal(wop_operator* op, int* id) {
    if (!op) return fail(-1, "bad argument");
    Property p;
    p.kind = P_METAL;
    return add_prop(op, p, id);
}

int wop_add_lorentz_material(wop_operator* op, double eps, double kappa, double mue, double sigma, double fp, double tr, double fl, int* id) {
    if (!op) return fail(-1, "bad argument");
    Property p;
    p.kind = P_MATERIAL; p.lorentz = true;
    p.eps = eps; p.kappa = kappa; p.mue = mue; p.sigma = sigma;
    p.eps_plasma_f = fp; p.eps_relax_t = tr; p.eps_lor_f = fl;
    return add_prop(op, p, id);
}

int wop_add_lumped_element(wop_operator* op, int ny, double R, double C, int caps, int* id) {
    if (!op || ny < 0 || ny > 2) return fail(-1, "bad argument");
    Property p;
    p.kind = P_LUMPED; p.ny = ny; p.R = R; p.C = C; p.caps = caps != 0;
    return add_prop(op, p, id);
}

int wop_add_lumped_rlc(wop_operator* op, int ny, int type, double R, double L, double C, int caps, int* id) {
    if (!op || ny < 0 || ny > 2 || type < 0 || type > 1) return fail(-1, "bad argument");
    Property p;
    p.kind = P_LUMPED; p.ny = ny; p.le_type = type; p.R = R; p.L = L; p.C = C; p.caps = caps != 0;
    return add_prop(op, p, id);
}

int wop_get_rlc_count(wop_operator* op, uint32_t* count) {
    if (!op || !count || !op->built) return fail(-1, "operator not built");
    *count = (uint32_t)op->rlc.dir.size();
    return 0;
}

int wop_get_rlc(wop_operator* op, uint32_t* pos, uint32_t* dir, int dtype, void* coef) {
    if (!op || !op->built || !pos || !dir || !coef) return fail(-1, "bad argument");
    const size_t n = op->rlc.dir.size();
    for (int k = 0; k < 3; ++k) std::copy(op->rlc.pos[k].begin(), op->rlc.pos[k].end(), pos + k * n);
    std::copy(op->rlc.dir.begin(), op->rlc.dir.end(), dir);
    for (int k = 0; k < 9; ++k)
        for (size_t i = 0; i < n; ++i) {
            if (dtype == 1) ((double*)coef)[k * n + i] = op->rlc.c[k][i];
            else ((float*)coef)[k * n + i] = (float)op->rlc.c[k][i];
        }
    return 0;
}

int wop_add_excitation(wop_operator* op, int type, const double value[3], double delay, int* id) {
    if (!op || !value || type < 0 || type > 3) return fail(-1, "bad argument");
    Property p;
    p.kind = P_EXCITATION; p.exc_type = type; p.delay = delay;
    for (int n = 0; n < 3; ++n) p.exc_val[n] = value[n];
    return add_prop(op, p, id);
}

int wop_set_plane_wave(wop_operator* op, const double start[3], const double stop[3], const double prop_dir[3],
                       const double e_vec[3], double freq) {
    if (!op || !start || !stop || !prop_dir || !e_vec) return fail(-1, "bad argument");
    for (int n = 0; n < 3; ++n) {
        op->pw.a[n] = start[n]; op->pw.b[n] = stop[n];
        op->pw.k[n] = prop_dir[n]; op->pw.E[n] = e_vec[n];
    }
    op->pw.f = freq;
    op->pw.set = true;
    return 0;
}

int wop_get_tfsf_info(wop_operator* op, uint32_t* volt_count, uint32_t* curr_count, uint32_t* max_delay) {
    if (!op || !volt_count || !curr_count || !max_delay || !op->built) return fail(-1, "operator not built");
    *volt_count = (uint32_t)op->tfsf_v.size();
    *curr_count = (uint32_t)op->tfsf_i.size();
    *max_delay = op->tfsf_max_delay;
    return 0;
}

int wop_get_tfsf(wop_operator* op, int which, uint32_t* pos, uint32_t* comp, uint32_t* delay, double* delta, double* amp) {
    if (!op || !op->built || (which != 0 && which != 1)) return fail(-1, "bad argument");
    const auto& L = which == 0 ? op->tfsf_v : op->tfsf_i;
    const size_t n = L.size();
    for (size_t i = 0; i < n; ++i) {
        const auto& e = L[i];
        if (pos) for (int a = 0; a < 3; ++a) pos[a * n + i] = e.pos[a];
        if (comp) comp[i] = e.comp;
        if (delay) delay[i] = e.delay;
        if (delta) delta[i] = e.delta;
        if (amp) amp[i] = e.amp;
    }
    return 0;
}

static int add_prim(wop_operator* op, Primitive pr) {
    // descending priority, first added wins among equals
    auto it = std::upper_bound(op->prims.begin(), op->prims.end(), pr,
                               [](const Primitive& a, const Primitive& b) { return a.priority > b.priority; });
    op->prims.insert(it, pr);
    op->xcand.clear();
    return 0;
}

int wop_add_box(wop_operator* op, int prop, const double start[3], const double stop[3], int priority) {
    if (!op || !start || !stop || prop < 0 || prop >= (int)op->props.size()) return fail(-1, "bad argument");
    Primitive pr;
    pr.shape = 0; pr.prop = prop; pr.priority = priority;
    for (int n = 0; n < 3; ++n) {
        pr.a[n] = std::min(start[n], stop[n]);
        pr.b[n] = std::max(start[n], stop[n]);
        pr.bb0[n] = pr.a[n]; pr.bb1[n] = pr.b[n];
    }
    return add_prim(op, pr);
}

int wop_add_sphere(wop_operator* op, int prop, const double center[3], double radius, int priority) {
    if (!op || !center || prop < 0 || prop >= (int)op->props.size() || !(radius > 0)) return fail(-1, "bad argument");
    Primitive pr;
    pr.shape = 1; pr.prop = prop; pr.priority = priority;
    for (int n = 0; n < 3; ++n) {
        pr.a[n] = center[n];
        pr.bb0[n] = center[n] - radius; pr.bb1[n] = center[n] + radius;
    }
    pr.b[0] = radius;
    return add_prim(op, pr);
}

int wop_set_excite_signal(wop_operator* op, int type, double f0, double fc) {
    if (!op || type < 0 || type > 3) return fail(-1, "bad argument");
    if (type == WOP_EXC_GAUSS && !(fc > 0)) return fail(-1, "Gaussian pulse needs fc > 0");
    if (type == WOP_EXC_SINE && !(f0 > 0)) return fail(-1, "sine needs f0 > 0");
    op->exc_sig_type = type; op->f0 = f0; op->fc = fc;
    return 0;
}

namespace {
// workers for the length of one build call
struct PoolScope {
    const wop_operator* op;
    explicit PoolScope(const wop_operator* o) : op(o) { ++op->pool_depth; }
    ~PoolScope() {
        if (--op->pool_depth == 0) op->pool.reset();
    }
};
}  // namespace

int wop_build(wop_operator* op, int prec, uint32_t nr_ts) {
    if (!op) return fail(-1, "bad argument");
    for (int n = 0; n < 3; ++n)
        if (op->N[n] < 2) return fail(-1, "mesh not set on axis %d", n);
    op->prec = prec;
    PoolScope scope(op);
    if (prec == 0) return op->build<float>(nr_ts);
    return op->build<double>(nr_ts);
}

int wop_get_num_lines(wop_operator* op, uint32_t n[3]) {
    if (!op || !n) return fail(-1, "bad argument");
    for (int k = 0; k < 3; ++k) n[k] = op->N[k];
    return 0;
}

int wop_get_edge_lengths(wop_operator* op, int dual, double* out) {
    if (!op || !out) return fail(-1, "bad argument");
    if (op->lines[0].empty() || op->lines[1].empty() || op->lines[2].empty()) return fail(-1, "mesh not set");
    size_t k = 0;
    for (int n = 0; n < 3; ++n)
        for (uint32_t i = 0; i < op->N[n]; ++i) {
            uint32_t pos[3] = {0, 0, 0};
            pos[n] = i;
            out[k++] = op->EdgeLength(n, pos, dual != 0);
        }
    return 0;
}

int wop_set_material_store_flags(wop_operator* op, int type, int val) {
    if (!op) return fail(-1, "bad argument");
    if (type < 0 || type > 2) return fail(-1, "material type must be 0 (epsilon), 1 (kappa) or 2 (mu)");
    op->store_mat[type] = val != 0;
    return 0;
}

int wop_get_material(wop_operator* op, int type, float* out) {
    if (!op || !out) return fail(-1, "bad argument");
    if (type < 0 || type > 2) return fail(-1, "material type must be 0 (epsilon), 1 (kappa) or 2 (mu)");
    if (!op->built) return fail(-1, "operator not built");
    if (op->mat[type][0].size() != op->nd()) return fail(-1, "material type %d was not stored by the last build", type);
    const size_t n = op->nd();
    for (int d = 0; d < 3; ++d) memcpy(out + d * n, op->mat[type][d].data(), n * sizeof(float));
    return 0;
}

int wop_get_timestep(wop_operator* op, double* dT) {
    if (!op || !dT || !op->built) return fail(-1, "operator not built");
    *dT = op->dT;
    return 0;
}

int wop_get_excite_info(wop_operator* op, uint32_t* length, uint32_t* nyquist, uint32_t* max_step) {
    if (!op || !op->built) return fail(-1, "operator not built");
    if (length) *length = op->exc_length;
    if (nyquist) *nyquist = op->nyquist;
    if (max_step) *max_step = op->max_step;
    return 0;
}

int wop_get_signal_period(wop_operator* op, double* period, uint32_t* period_ts) {
    if (!op || !op->built) return fail(-1, "operator not built");
    // Excitation::SetupSinusoidal sets the period 1/f0, every other signal 0
    // (FDTD/excitation.cpp:55-93); the excitation hooks wrap the signal index
    // at int(period/dT) (engine_ext_excitation.cpp:43-45, engine_ext_tfsf.cpp:41)
    const double T = (op->exc_sig_type == WOP_EXC_SINE && op->f0 > 0) ? 1 / op->f0 : 0.0;
    if (period) *period = T;
    if (period_ts) *period_ts = T > 0 ? (uint32_t)(int)(T / op->dT) : 0;
    return 0;
}

int wop_get_signal(wop_operator* op, double* v, double* i) {
    if (!op || !op->built || !v || !i) return fail(-1, "operator not built");
    memcpy(v, op->sig_v.data(), op->sig_v.size() * 8);
    memcpy(i, op->sig_i.data(), op->sig_i.size() * 8);
    return 0;
}

static void put(const std::vector<double>& src, int dtype, void* dst) {
    if (!dst) return;
    if (dtype == 1) memcpy(dst, src.data(), src.size() * 8);
    else {
        float* o = (float*)dst;
        for (size_t i = 0; i < src.size(); ++i) o[i] = (float)src[i];
    }
}
static void put_f(const std::vector<float>& src, int dtype, void* dst) {
    if (!dst) return;
    if (dtype == 1) {
        double* o = (double*)dst;
        for (size_t i = 0; i < src.size(); ++i) o[i] = (double)src[i];
    } else memcpy(dst, src.data(), src.size() * 4);
}

int wop_get_coeffs(wop_operator* op, int dtype, void* vv, void* vi, void* ii, void* iv) {
    if (!op || !op->built) return fail(-1, "operator not built");
    if (op->streamed) return fail(-1, "the streamed build keeps no full coefficient arrays");
    void* dst[4] = {vv, vi, ii, iv};
    for (int k = 0; k < 4; ++k) {
        if (op->prec == 0) put3(op->sf.co[k], dtype, dst[k]);
        else put3(op->sd.co[k], dtype, dst[k]);
    }
    return 0;
}

int wop_get_ec(wop_operator* op, double* C, double* G, double* L, double* R) {
    if (!op || !op->built) return fail(-1, "operator not built");
    if (op->streamed) return fail(-1, "the streamed build keeps no EC arrays");
    void* dst[4] = {C, G, L, R};
    for (int k = 0; k < 4; ++k) {
        if (op->prec == 0) put3(op->sf.EC[k], 1, dst[k]);
        else put3(op->sd.EC[k], 1, dst[k]);
    }
    return 0;
}

}  // extern "C"

namespace {
struct TupleHash {
    template <typename T, size_t K>
    size_t operator()(const std::array<T, K>& t) const {
        uint64_t h = 1469598103934665603ull;
        for (T v : t) { h ^= (uint64_t)v; h *= 1099511628211ull; h ^= h >> 29; }
        return (size_t)h;
    }
};

// Deduplicates the tuples of one x plane into per-cell ids of a table that
// grows over the planes, numbering tuples in cell order of first occurrence.
// Rows are deduplicated in parallel (local ids in row order of first
// occurrence; runs along z and short rows need no hashing), then merged row
// by row into the table, then the row ids are mapped to table ids.
template <typename Tuple>
struct PlaneDedup {
    struct Row {
        std::vector<Tuple> uniq;
        // open-addressing index into uniq (slot = id or UINT32_MAX), used
        // once a row holds more than 16 tuples
        std::vector<uint32_t> slots;
        std::vector<uint32_t> global;
        uint32_t find(const Tuple& t, size_t h) const {
            const size_t mask = slots.size() - 1;
            for (size_t k = h & mask;; k = (k + 1) & mask) {
                if (slots[k] == UINT32_MAX) return UINT32_MAX;
                if (uniq[slots[k]] == t) return slots[k];
            }
        }
        void insert(uint32_t id, size_t h) {
            const size_t mask = slots.size() - 1;
            size_t k = h & mask;
            while (slots[k] != UINT32_MAX) k = (k + 1) & mask;
            slots[k] = id;
        }
        void reindex() {
            size_t n = 64;                  // a power of two, at least 4x the tuples
            while (n < 4 * uniq.size()) n *= 2;
            slots.assign(n, UINT32_MAX);
            TupleHash H;
            for (uint32_t q = 0; q < uniq.size(); ++q) insert(q, H(uniq[q]));
        }
    };
    std::vector<Row> rows;
    std::unordered_map<Tuple, uint32_t, TupleHash> map;
    std::vector<Tuple> uniq;

    // get(y, z, tuple&) fills the tuple of cell (y, z); out has Ny*Nz ids.
    // same(y, z) may say that cell (y, z) holds the tuple of (y, z-1), which
    // is then not built.  Returns false once the table holds more than
    // max_entries (0 = no limit).
    template <typename Get>
    bool plane(const wop_operator* op, uint32_t* out, size_t max_entries, Get get) {
        return plane(op, out, max_entries, get, [](uint32_t, uint32_t) { return false; });
    }
    template <typename Get, typename Same>
    bool plane(const wop_operator* op, uint32_t* out, size_t max_entries, Get get, Same same) {
        const uint32_t Ny = op->N[1], Nz = op->N[2];
        rows.resize(Ny);
        op->parallel_x(Ny, [&](uint32_t y) {
            Row& r = rows[y];
            r.uniq.clear();
            uint32_t* o = out + (size_t)y * Nz;
            uint32_t prev = UINT32_MAX;
            Tuple t;
            TupleHash H;
            for (uint32_t z = 0; z < Nz; ++z) {
                if (prev != UINT32_MAX && same(y, z)) {
                    o[z] = prev;
                    continue;
                }
                get(y, z, t);
                uint32_t id = UINT32_MAX;
                size_t h = 0;
                if (prev != UINT32_MAX && t == r.uniq[prev]) id = prev;
                else if (r.uniq.size() <= 16) {
                    for (uint32_t q = 0; q < r.uniq.size(); ++q)
                        if (r.uniq[q] == t) { id = q; break; }
                } else {
                    h = H(t);
                    id = r.find(t, h);
                }
                if (id == UINT32_MAX) {
                    id = (uint32_t)r.uniq.size();
                    r.uniq.push_back(t);
                    if (r.uniq.size() == 17) r.reindex();
                    else if (r.uniq.size() > 17) {
                        if (2 * r.uniq.size() > r.slots.size()) r.reindex();
                        else r.insert(id, h);
                    }
                }
                o[z] = id;
                prev = id;
            }
        });
        // tuples the table held before this plane: looked up in parallel
        // (the table is not modified meanwhile); the rest, in row order
        op->parallel_x(Ny, [&](uint32_t y) {
            Row& r = rows[y];
            r.global.assign(r.uniq.size(), UINT32_MAX);
            for (size_t q = 0; q < r.uniq.size(); ++q) {
                auto it = map.find(r.uniq[q]);
                if (it != map.end()) r.global[q] = it->second;
            }
        });
        for (uint32_t y = 0; y < Ny; ++y) {
            Row& r = rows[y];
            for (size_t q = 0; q < r.uniq.size(); ++q) {
                if (r.global[q] != UINT32_MAX) continue;
                auto it = map.find(r.uniq[q]);
                if (it == map.end()) {
                    const uint32_t id = (uint32_t)uniq.size();
                    map.emplace(r.uniq[q], id);
                    uniq.push_back(r.uniq[q]);
                    r.global[q] = id;
                } else r.global[q] = it->second;
            }
            if (max_entries && uniq.size() > max_entries) return false;
        }
        op->parallel_x(Ny, [&](uint32_t y) {
            const Row& r = rows[y];
            uint32_t* o = out + (size_t)y * Nz;
            for (uint32_t z = 0; z < Nz; ++z) o[z] = r.global[o[z]];
        });
        return true;
    }
};

// Float build that never holds full per-cell arrays: x planes pass through a
// ring of three plane slots.  Pass 1 computes EC plane by plane and the
// Var3 timestep from planes x-1..x+1, and keeps each EC plane as per-cell ids
// into a table of distinct EC 12-tuples; pass 2 restores EC from those ids,
// computes the coefficients, boundary conditions, PEC, lumped elements,
// excitation and UPML of each plane in the full build's order (every step
// touches only its own cell) and deduplicates the plane's coefficient
// 12-tuples into index[] in cell order -- the same index and table the full
// build followed by wop_compress gives.
int stream_build(wop_operator* op, uint32_t nr_ts, uint32_t* index, uint32_t max_entries, uint32_t* n_entries) {
    if (op->pw.set) return fail(-1, "the streamed build has no TFSF plane wave");
    for (auto& p : op->props) {
        if (p.lorentz) return fail(-1, "the streamed build has no Lorentz/Drude materials");
        if ((p.kind & P_LUMPED) && ((p.le_type == 0 && p.L > 0.0) || p.le_type == 1))
            return fail(-1, "the streamed build has no lumped RLC elements");
    }
    op->rlc = wop_operator::Rlc();
    const uint32_t* N = op->N;
    op->sf.clear();
    op->sd.clear();
    op->FillEdgeTables();
    op->InitMaterialStorage();
    op->built = false;
    op->prec = 0;
    op->ring_slot.resize(N[0]);
    for (uint32_t x = 0; x < N[0]; ++x) op->ring_slot[x] = x % 3;
    op->ring = 3;
    op->streamed = true;
    op->stream_table.clear();
    const size_t plane = (size_t)N[1] * N[2];
    Builder<float> b(op);
    for (int k = 0; k < 4; ++k)
        for (int d = 0; d < 3; ++d) { b.EC[k][d].assign(3 * plane, 0.0f); b.co[k][d].assign(3 * plane, 0.0f); }
    b.same_prev.assign(3 * plane, 0);
    // EC planes kept as ids (4 bytes per cell) into their distinct tuples
    std::vector<uint32_t> ec_ids((size_t)N[0] * plane);
    PlaneDedup<std::array<uint32_t, 12>> ec_tab;
    auto keep_ec = [&](uint32_t x) {
        ec_tab.plane(op, ec_ids.data() + (size_t)x * plane, 0, [&](uint32_t y, uint32_t z, std::array<uint32_t, 12>& t) {
            const size_t i = op->idx(x, y, z);
            for (int k = 0; k < 4; ++k)
                for (int d = 0; d < 3; ++d) memcpy(&t[k * 3 + d], &b.EC[k][d][i], 4);
        }, [&](uint32_t y, uint32_t z) { return b.same_prev[op->idx(x, y, z)] != 0; });
    };
    // pass 1: timestep
    double dT = 1e200;
    b.ECPlane(0);
    keep_ec(0);
    if (N[0] > 1) { b.ECPlane(1); keep_ec(1); }
    for (uint32_t x = 0; x < N[0]; ++x) {
        if (x >= 1 && x + 1 < N[0]) { b.ECPlane(x + 1); keep_ec(x + 1); }
        const double t = b.TimestepPlane(x);
        if (t < dT && t > 0.0) dT = t;
        if (op->verbose && ((x + 1) % 128 == 0 || x + 1 == N[0]))
            fprintf(stderr, "streamed build: timestep pass, plane %u of %u, %zu distinct EC tuples\n", x + 1, N[0], ec_tab.uniq.size());
    }
    ec_tab.map.clear();
    ec_tab.rows.clear();
    if (op->forced_dT > 0) op->dT = op->forced_dT;
    else op->dT = dT;
    op->dT *= op->ts_factor;
    if (op->exc_sig_type == WOP_EXC_SINE && op->f0 > 0) {
        double period = 1 / op->f0;
        unsigned TS = (unsigned)ceil(period / op->dT);
        op->dT = period / TS;
    }
    // pass 2: coefficients plane by plane, compressed as they are finished
    b.UpmlSetup();
    op->exc_v.clear();
    op->exc_i.clear();
    op->tfsf_v.clear();
    op->tfsf_i.clear();
    op->tfsf_max_delay = 0;
    op->ade.clear();
    PlaneDedup<std::array<uint64_t, 12>> co_tab;
    for (uint32_t x = 0; x < N[0]; ++x) {
        // EC of plane x as pass 1 computed it
        op->parallel_x(N[1], [&](uint32_t y) {
            const uint32_t* id = ec_ids.data() + (size_t)x * plane + (size_t)y * N[2];
            for (uint32_t z = 0; z < N[2]; ++z) {
                const size_t i = op->idx(x, y, z);
                const std::array<uint32_t, 12>& t = ec_tab.uniq[id[z]];
                for (int k = 0; k < 4; ++k)
                    for (int d = 0; d < 3; ++d) memcpy(&b.EC[k][d][i], &t[k * 3 + d], 4);
            }
        });
        op->parallel_x(N[1], [&](uint32_t y) {
            // the coefficients are a function of the cell's EC tuple and dT alone
            const uint32_t* id = ec_ids.data() + (size_t)x * plane + (size_t)y * N[2];
            for (uint32_t z = 0; z < N[2]; ++z) {
                const size_t i = op->idx(x, y, z);
                if (op->row_reuse && z > 0 && id[z] == id[z - 1]) {
                    for (int k = 0; k < 4; ++k)
                        for (int d = 0; d < 3; ++d) b.co[k][d][i] = b.co[k][d][i - 1];
                    continue;
                }
                for (int d = 0; d < 3; ++d) b.ECOperatorPos(d, i);
            }
        });
        b.ApplyElectricBC((int)x);
        b.PECPlane(x);
        b.LumpedElements((int)x);
        b.ApplyMagneticBC((int)x);
        b.Excitation((int)x);
        b.UpmlPlane(x);
        const bool ok = co_tab.plane(op, index + (size_t)x * plane, max_entries,
                                     [&](uint32_t y, uint32_t z, std::array<uint64_t, 12>& t) {
            const size_t i = op->idx(x, y, z);
            for (int k = 0; k < 4; ++k)
                for (int d = 0; d < 3; ++d) {
                    const double v = (double)b.co[k][d][i];
                    memcpy(&t[k * 3 + d], &v, 8);
                }
        }, [&](uint32_t y, uint32_t z) {
            // the tuple widens the 12 floats, so equal float bits give equal tuples
            const size_t i = op->idx(x, y, z);
            for (int k = 0; k < 4; ++k)
                for (int d = 0; d < 3; ++d)
                    if (memcmp(&b.co[k][d][i], &b.co[k][d][i - 1], sizeof(float))) return false;
            return true;
        });
        if (!ok) {
            op->ring = 0;
            *n_entries = (uint32_t)co_tab.uniq.size();
            return fail(-2, "more than %u unique coefficient tuples", max_entries);
        }
        if (op->verbose && ((x + 1) % 128 == 0 || x + 1 == N[0]))
            fprintf(stderr, "streamed build: coefficient pass, plane %u of %u, %zu unique tuples\n", x + 1, N[0], co_tab.uniq.size());
    }
    std::vector<uint32_t>().swap(ec_ids);
    std::vector<std::array<uint64_t, 12>>& uniq = co_tab.uniq;
    b.UpmlFinish();
    // excitation lists in the full build's z, y, x order
    auto zyx = [](const wop_operator::Exc& a, const wop_operator::Exc& c) {
        if (a.pos[2] != c.pos[2]) return a.pos[2] < c.pos[2];
        if (a.pos[1] != c.pos[1]) return a.pos[1] < c.pos[1];
        return a.pos[0] < c.pos[0];
    };
    std::stable_sort(op->exc_v.begin(), op->exc_v.end(), zyx);
    std::stable_sort(op->exc_i.begin(), op->exc_i.end(), zyx);
    b.Signal(nr_ts);
    b.Mur();
    b.EC[0][0].clear();
    for (int k = 0; k < 4; ++k)
        for (int d = 0; d < 3; ++d) { std::vector<float>().swap(b.EC[k][d]); std::vector<float>().swap(b.co[k][d]); }
    op->stream_table.resize(uniq.size() * 12);
    for (size_t r = 0; r < uniq.size(); ++r)
        for (int k = 0; k < 12; ++k) memcpy(&op->stream_table[r * 12 + k], &uniq[r][k], 8);
    *n_entries = (uint32_t)uniq.size();
    op->ring = 0;
    op->built = true;
    return 0;
}
}  // namespace

extern "C" int wop_build_streamed(wop_operator* op, uint32_t nr_ts, uint32_t* index, uint32_t max_entries, uint32_t* n_entries) {
    if (!op || !index || !n_entries) return fail(-1, "bad argument");
    if (op->lines[0].empty() || op->lines[1].empty() || op->lines[2].empty()) return fail(-1, "mesh not set");
    PoolScope scope(op);
    try {
        return stream_build(op, nr_ts, index, max_entries, n_entries);
    } catch (const std::bad_alloc&) {
        op->ring = 0;
        return fail(-3, "out of host memory");
    }
}

extern "C" int wop_get_streamed_table(wop_operator* op, double* table) {
    if (!op || !table || !op->built || !op->streamed) return fail(-1, "no streamed build");
    memcpy(table, op->stream_table.data(), op->stream_table.size() * sizeof(double));
    return 0;
}

extern "C" int wop_compress(wop_operator* op, uint32_t* n_entries, uint32_t* index, double* table, uint32_t max_entries) {
    if (!op || !op->built || !n_entries) return fail(-1, "operator not built");
    if (op->streamed) return fail(-1, "the streamed build is compressed already (wop_get_streamed_table)");
    const size_t n = op->nd();
    std::unordered_map<std::array<uint64_t, 12>, uint32_t, TupleHash> map;
    std::vector<std::array<uint64_t, 12>> uniq;
    for (size_t i = 0; i < n; ++i) {
        std::array<uint64_t, 12> t;
        for (int k = 0; k < 4; ++k)
            for (int d = 0; d < 3; ++d) {
                double v = op->coef_at(k, d, i);
                uint64_t bits;
                memcpy(&bits, &v, 8);
                t[k * 3 + d] = bits;
            }
        auto it = map.find(t);
        uint32_t id;
        if (it == map.end()) {
            id = (uint32_t)uniq.size();
            map.emplace(t, id);
            uniq.push_back(t);
            if (max_entries && uniq.size() > max_entries) {
                *n_entries = (uint32_t)uniq.size();
                return fail(-2, "more than %u unique coefficient tuples", max_entries);
            }
        } else id = it->second;
        if (index) index[i] = id;
    }
    *n_entries = (uint32_t)uniq.size();
    if (table)
        for (size_t r = 0; r < uniq.size(); ++r)
            for (int k = 0; k < 12; ++k) memcpy(&table[r * 12 + k], &uniq[r][k], 8);
    return 0;
}

// m rounded to `bits` significant bits (ties away from zero)
static float round_mantissa(double m, int bits) {
    if (m == 0.0) return 0.0f;
    int e;
    const double f = std::frexp(m, &e);
    return (float)std::ldexp(std::round(std::ldexp(f, bits)), e - bits);
}

// Factored coefficients: the EC formulas (Calc_ECOperatorPos,
// FDTD/operator.cpp:943-971, with C = eps*A/L and L = mue*A'/L' from
// Calc_ECPos, :1176-1224) make vi a material term times dT*L/A, L the primary
// edge along n and A the dual area across it, and iv a material term times
// dT*L'/A', L' the dual edge and A' the primary area.  Both mesh terms are
// products of one factor per axis, so with float factors
//   fv[n][axis][i] = P_n(i) (axis n) or 1/D_axis(i), fc[n][axis][i] = D_n(i) or 1/P_axis(i)
// (the x factor carrying dT/EPS0 resp. dT/MUE0) the coefficients are
//   vi = m_E * ((fv[n][0][x] * fv[n][1][y]) * fv[n][2][z]),  iv likewise,
// in float arithmetic in that order, with m = vi / g rounded to `bits`
// significant bits.  Cells of one material share m however the mesh is
// graded, so the tuples {vv, m_E, ii, m_H} collapse to far fewer rows than
// {vv, vi, ii, iv}.  The UPML rows and lumped elements are ordinary tuples.
extern "C" int wop_factor(wop_operator* op, int bits, uint32_t max_entries, uint32_t* index, float* table, uint32_t* n_entries,
                          float* fv, float* fc, double* max_rel_err) {
    if (!op || !op->built || !n_entries || !index || !table || !fv || !fc || !max_entries) return fail(-1, "bad argument");
    if (op->streamed) return fail(-1, "the streamed build keeps no per-cell coefficients");
    if (bits < 8 || bits > 23) return fail(-1, "bits must be 8..23");
    const uint32_t Nx = op->N[0], Ny = op->N[1], Nz = op->N[2];
    const size_t L = (size_t)Nx + Ny + Nz;
    for (int n = 0; n < 3; ++n) {
        size_t k = 0;
        for (int a = 0; a < 3; ++a)
            for (uint32_t i = 0; i < op->N[a]; ++i, ++k) {
                uint32_t pos[3] = {0, 0, 0};
                pos[a] = i;
                const double P = op->EdgeLength(a, pos, false), D = op->EdgeLength(a, pos, true);
                if (!(P > 0) || !(D > 0)) return fail(-1, "edge length <= 0 on axis %d line %u", a, i);
                double e = a == n ? P : 1.0 / D;
                double h = a == n ? D : 1.0 / P;
                if (a == 0) { e *= op->dT / EPS0; h *= op->dT / MUE0; }
                fv[n * L + k] = (float)e;
                fc[n * L + k] = (float)h;
                if (!std::isfinite(fv[n * L + k]) || fv[n * L + k] == 0.0f || !std::isfinite(fc[n * L + k]) || fc[n * L + k] == 0.0f)
                    return fail(-1, "mesh factor out of float range on axis %d line %u", a, i);
            }
    }
    std::vector<double> rowerr(Ny, 0.0);
    PlaneDedup<std::array<uint32_t, 12>> tab;
    const size_t plane = (size_t)Ny * Nz;
    for (uint32_t x = 0; x < Nx; ++x) {
        const bool ok = tab.plane(op, index + (size_t)x * plane, max_entries, [&](uint32_t y, uint32_t z, std::array<uint32_t, 12>& t) {
            const size_t i = op->idx(x, y, z);
            const uint32_t at[3] = {x, Nx + y, Nx + Ny + z};
            for (int n = 0; n < 3; ++n) {
                float q[4];
                q[0] = (float)op->coef_at(0, n, i);
                q[2] = (float)op->coef_at(2, n, i);
                for (int side = 0; side < 2; ++side) {
                    const float* f = (side ? fc : fv) + n * L;
                    const float g = (f[at[0]] * f[at[1]]) * f[at[2]];
                    const float c = (float)op->coef_at(side ? 3 : 1, n, i);
                    const float m = round_mantissa((double)c / (double)g, bits);
                    const float r = m * g;
                    if (c != 0.0f) rowerr[y] = std::max(rowerr[y], std::fabs((double)r - c) / std::fabs((double)c));
                    else if (r != 0.0f) rowerr[y] = INFINITY;
                    q[side ? 3 : 1] = m;
                }
                for (int k = 0; k < 4; ++k) memcpy(&t[k * 3 + n], &q[k], 4);
            }
        });
        if (!ok) {
            *n_entries = (uint32_t)tab.uniq.size();
            return fail(-2, "more than %u unique factored tuples", max_entries);
        }
    }
    *n_entries = (uint32_t)tab.uniq.size();
    for (size_t r = 0; r < tab.uniq.size(); ++r)
        for (int k = 0; k < 12; ++k) memcpy(&table[r * 12 + k], &tab.uniq[r][k], 4);
    if (max_rel_err) *max_rel_err = *std::max_element(rowerr.begin(), rowerr.end());
    return 0;
}

extern "C" {

int wop_get_excitation_counts(wop_operator* op, uint32_t* vc, uint32_t* cc) {
    if (!op || !op->built) return fail(-1, "operator not built");
    if (vc) *vc = (uint32_t)op->exc_v.size();
    if (cc) *cc = (uint32_t)op->exc_i.size();
    return 0;
}

int wop_get_excitation(wop_operator* op, uint32_t* vpos, uint32_t* vdir, double* vamp, uint32_t* vdel,
                       uint32_t* cpos, uint32_t* cdir, double* camp, uint32_t* cdel) {
    if (!op || !op->built) return fail(-1, "operator not built");
    auto fill = [](const std::vector<wop_operator::Exc>& l, uint32_t* pos, uint32_t* dir, double* amp, uint32_t* del) {
        const size_t c = l.size();
        for (size_t i = 0; i < c; ++i) {
            if (pos) for (int n = 0; n < 3; ++n) pos[n * c + i] = l[i].pos[n];
            if (dir) dir[i] = l[i].dir;
            if (amp) amp[i] = l[i].amp;
            if (del) del[i] = l[i].delay;
        }
    };
    fill(op->exc_v, vpos, vdir, vamp, vdel);
    fill(op->exc_i, cpos, cdir, camp, cdel);
    return 0;
}

int wop_get_num_upml(wop_operator* op, uint32_t* n) {
    if (!op || !op->built || !n) return fail(-1, "operator not built");
    *n = (uint32_t)op->upml.size();
    return 0;
}

int wop_get_upml_range(wop_operator* op, uint32_t k, uint32_t start[3], uint32_t num_lines[3]) {
    if (!op || !op->built || k >= op->upml.size()) return fail(-1, "bad UPML index");
    for (int n = 0; n < 3; ++n) { start[n] = op->upml[k].start[n]; num_lines[n] = op->upml[k].n[n]; }
    return 0;
}

int wop_get_upml(wop_operator* op, uint32_t k, int dtype, void* vv, void* vvfo, void* vvfn, void* ii, void* iifo, void* iifn) {
    if (!op || !op->built || k >= op->upml.size()) return fail(-1, "bad UPML index");
    void* dst[6] = {vv, vvfo, vvfn, ii, iifo, iifn};
    for (int j = 0; j < 6; ++j) {
        const wop_operator::Upml& u = op->upml[k];
        if (!u.af[j].empty() || u.a[j].empty()) put_f(u.af[j], dtype, dst[j]);
        else put(u.a[j], dtype, dst[j]);
    }
    return 0;
}

int wop_get_num_ade(wop_operator* op, uint32_t* orders) {
    if (!op || !op->built || !orders) return fail(-1, "operator not built");
    *orders = (uint32_t)op->ade.size();
    return 0;
}

int wop_get_ade_info(wop_operator* op, uint32_t o, uint32_t* count, uint32_t* flags) {
    if (!op || !op->built || o >= op->ade.size()) return fail(-1, "bad ADE order");
    if (count) *count = (uint32_t)op->ade[o].pos[0].size();
    if (flags) *flags = op->ade[o].flags;
    return 0;
}

int wop_get_ade(wop_operator* op, uint32_t o, uint32_t* pos, int dtype, void* v_int, void* v_ext, void* v_lor,
                void* i_int, void* i_ext, void* i_lor) {
    if (!op || !op->built || o >= op->ade.size()) return fail(-1, "bad ADE order");
    const auto& a = op->ade[o];
    const size_t c = a.pos[0].size();
    if (pos) for (int n = 0; n < 3; ++n) memcpy(pos + n * c, a.pos[n].data(), c * 4);
    void* dst[6] = {v_int, v_ext, v_lor, i_int, i_ext, i_lor};
    for (int j = 0; j < 6; ++j) put(a.lists[j], dtype, dst[j]);
    return 0;
}

int wop_snap_to_mesh_line(wop_operator* op, int axis, double coord, int dual, uint32_t* idx, int* inside) {
    if (!op || axis < 0 || axis > 2 || !idx || op->N[axis] < 2) return fail(-1, "bad argument");
    bool in = false;
    *idx = op->SnapToMeshLine(axis, coord, in, dual != 0);
    if (inside) *inside = in ? 1 : 0;
    return 0;
}

}  // extern "C"
