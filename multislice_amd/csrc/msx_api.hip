// msx_api.hip -- the C ABI of include/msx.h: context, device tables, and the
// host-side orchestration of the sweep kernels for each reference entry point.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/msx.h"
#include "msx_internal.h"
#include "msx_philox.h"
#include "fparams_tables.inc"
#include "doyle_turner.inc"

using namespace msx;

namespace msx {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

}  // namespace msx

namespace {

#define HIPCHK(x)                                                                                    \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess)                                                                        \
            return fail(e_ == hipErrorOutOfMemory ? MSX_ENOMEM : MSX_EHIP, "%s: %s (%s:%d)", #x,     \
                        hipGetErrorString(e_), __FILE__, __LINE__);                                  \
    } while (0)

#define CHK(x)                   \
    do {                         \
        int r_ = (x);            \
        if (r_ != MSX_OK) return r_; \
    } while (0)

struct DevBuf {
    void* p = nullptr;
    size_t n = 0;
    int ensure(size_t bytes) {
        if (bytes <= n) return MSX_OK;
        if (p) hipFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            p = nullptr;
            return fail(MSX_ENOMEM, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
        }
        n = bytes;
        return MSX_OK;
    }
    void release() {
        if (p) hipFree(p);
        p = nullptr;
        n = 0;
    }
};

// radix-independent twiddle table exp(+2 pi i m / n), m < n, from FP64.
// FP32 entries: among the float neighbours of the rounded (cos, sin) the pair
// whose c^2 + s^2 is closest to 1 (then the nearest to the exact value).
// Rounding cos and sin separately leaves |w|^2 - 1 of up to ~1e-7 with a
// fixed sign per entry (and a mean of -1.6e-8 over the 64th roots that the
// middle pass of a 512-point transform uses); the slice loop applies the same
// twiddles to the same strong beams every slice, so the wave's norm would
// drift by ~1e-5 over 100 slices.  The chosen pairs stay within 1.1e-7 rad
// of the exact angle.
template <typename C> std::vector<C> make_twiddles(int n) {
    using Re = decltype(C{}.x);
    std::vector<C> w(n);
    for (int m = 0; m < n; m++) {
        const long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)m / (long double)n;
        const double c = std::cos((double)a), s = std::sin((double)a);
        w[m].x = (Re)c;
        w[m].y = (Re)s;
        if (sizeof(Re) == 4) {
            double best_n = 1e30, best_d = 1e30;
            for (int dc = -1; dc <= 1; dc++)
                for (int ds = -1; ds <= 1; ds++) {
                    float fc = (float)c, fs = (float)s;
                    if (dc) fc = std::nextafter(fc, dc > 0 ? 2.0f : -2.0f);
                    if (ds) fs = std::nextafter(fs, ds > 0 ? 2.0f : -2.0f);
                    const double nerr = std::fabs((double)fc * fc + (double)fs * fs - 1.0);
                    const double derr = std::hypot((double)fc - c, (double)fs - s);
                    if (nerr < best_n || (nerr == best_n && derr < best_d)) {
                        best_n = nerr;
                        best_d = derr;
                        w[m].x = (Re)fc;
                        w[m].y = (Re)fs;
                    }
                }
        }
    }
    return w;
}

// freqn (slicelib.cpp:433-450) float k^2 table
std::vector<float> k2_table(int n, double a) {
    std::vector<float> k2(n);
    int imid = (int)(n / 2.0 + 0.5);
    for (int i = 0; i < n; i++) {
        float k = (i > imid) ? ((float)(i - n)) / ((float)a) : ((float)i) / ((float)a);
        k2[i] = k * k;
    }
    return k2;
}
// An oblique cell: gam set and not a right angle (to 1e-12 rad).
bool is_oblique(const msx_stack* st) { return st->gam != 0.0 && std::fabs(st->gam - M_PI / 2) > 1e-12; }

int fidx(int i, int n) {
    int imid = (int)(n / 2.0 + 0.5);
    return i > imid ? i - n : i;
}
// fraction of y columns inside the band (host copy of the kernels' test)
double band_fraction(const float* ky2_host, int ny, float k2max) {
    int k = 0;
    for (int i = 0; i < ny; i++) k += ky2_host[i] < k2max;
    return (double)k / ny;
}
// mulslice.cpp:481-489 (float arithmetic)
float k2max_of(int nx, int ny, double ax, double by, double bw) {
    float k2max = nx / (2.0F * (float)ax);
    float t = ny / (2.0F * (float)by);
    if (t < k2max) k2max = t;
    k2max = (float)bw * k2max;
    return k2max * k2max;
}

}  // namespace

// The fixed device array of msx_multislice_noalloc: one copy per GPU, held
// until the process ends.  Host code never uses this name as an address (in
// host code it is only a stand-in for the symbol): scratch_address() looks
// up each GPU's copy once, with that GPU current, and the entry point uses
// that address.
static __device__ __align__(256) unsigned char scratch[MSX_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device` (which must be the current device).
int scratch_address(int device, void** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0 || device >= kMaxDevices) return fail(MSX_EINVAL, "device %d: at most %d devices", device, kMaxDevices);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(MSX_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return MSX_OK;
}
}  // namespace

struct msx_ctx_s {
    int device = 0;
    void* scratch = nullptr;  // this GPU's copy of `scratch` (msx_multislice_noalloc)
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    // kernel profiling (msx_profile): per kernel kind launch count and time
    bool prof = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t ev_order = nullptr;  // msx_wait_stream
    long long prof_n[MSX_PROF_KINDS] = {0};
    double prof_ms[MSX_PROF_KINDS] = {0};
    double prof_bytes[MSX_PROF_KINDS] = {0};
    double band = 1.0;  // in-band fraction of y columns of the current slice loop

    // Launch through `go`; with profiling on, bracket it with events on the
    // launch stream and wait for it (serialises the stream: diagnostics only).
    template <typename F> int timed(int kind, double bytes, F&& go) {
        if (!prof) {
            HIPCHK(go());
            return MSX_OK;
        }
        HIPCHK(hipEventRecord(ev0, stream));
        HIPCHK(go());
        HIPCHK(hipEventRecord(ev1, stream));
        HIPCHK(hipEventSynchronize(ev1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
        prof_n[kind]++;
        prof_ms[kind] += ms;
        prof_bytes[kind] += bytes;
        return MSX_OK;
    }
    std::map<int, DevBuf> tw32, tw64;
    DevBuf wave, wave2, trans, aux, tabs, partial, outbuf, atoms, cells, vtmp, curvebuf, radtab, radcoef, binbuf,
        series;
    // host side of the potential's atom binning, kept across calls (no page
    // faults on reuse); the sorted records go to pinned memory for the upload
    std::vector<long long> h_cell;
    std::vector<AtomRec> h_rec;
    std::vector<int> h_cnt, h_fill;
    AtomRec* h_pin = nullptr;
    size_t h_pin_cap = 0;
    int ensure_pinned(size_t n) {
        if (n <= h_pin_cap) return MSX_OK;
        if (h_pin) HIPCHK(hipHostFree(h_pin));
        h_pin = nullptr;
        h_pin_cap = 0;
        HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&h_pin), sizeof(AtomRec) * n, hipHostMallocDefault));
        h_pin_cap = n;
        return MSX_OK;
    }

    // extra streams for job groups that run concurrently (msx_multislice_batch),
    // each with its own wave, pointer-table and propagator-table buffers
    struct Lane {
        hipStream_t stream = nullptr;
        hipEvent_t done = nullptr;
        DevBuf wave, aux, tabs;
    };
    std::vector<Lane> lanes;
    hipEvent_t lanes_go = nullptr;
    int ensure_lanes(int n) {
        if (!lanes_go) HIPCHK(hipEventCreateWithFlags(&lanes_go, hipEventDisableTiming));
        while ((int)lanes.size() < n) {
            Lane l;
            HIPCHK(hipStreamCreateWithFlags(&l.stream, hipStreamNonBlocking));
            HIPCHK(hipEventCreateWithFlags(&l.done, hipEventDisableTiming));
            lanes.push_back(std::move(l));
        }
        return MSX_OK;
    }
    // make lane i the context's stream and buffers (and back again)
    void swap_lane(int i) {
        Lane& l = lanes[i];
        std::swap(stream, l.stream);
        std::swap(wave, l.wave);
        std::swap(aux, l.aux);
        std::swap(tabs, l.tabs);
    }

    // launch shapes picked by msx_autotune / msx_set_tuning, per (nx, ny, prec)
    std::map<std::tuple<int, int, int>, msx_tuning> tune;
    const msx_tuning* tuning(int nx, int ny, int prec) const {
        auto it = tune.find(std::make_tuple(nx, ny, prec));
        return it == tune.end() ? nullptr : &it->second;
    }

    // hipGraph replay of slice loops (msx_set_graphs): a keying pass records
    // the kind and argument bytes of every launch instead of launching; the
    // key selects a cached executable graph captured from the same sequence.
    bool graphs = false;
    bool keying = false;
    std::string gkey;
    std::map<std::string, hipGraphExec_t> gcache;
    long long graph_hits = 0, graph_captures = 0;
    void key_add(int kind, int n, int mode, const void* args, size_t bytes) {
        const int hdr[3] = {kind, n, mode};
        gkey.append(reinterpret_cast<const char*>(hdr), sizeof hdr);
        gkey.append(static_cast<const char*>(args), bytes);
    }
    void drop_graphs() {
        for (auto& kv : gcache) hipGraphExecDestroy(kv.second);
        gcache.clear();
    }

    template <typename C> int twiddles(int n, const C** out) {
        auto& m = (sizeof(C) == 8) ? tw32 : tw64;
        DevBuf& b = m[n];
        if (!b.p) {
            std::vector<C> w = make_twiddles<C>(n);
            CHK(b.ensure(sizeof(C) * n));
            HIPCHK(hipMemcpy(b.p, w.data(), sizeof(C) * n, hipMemcpyHostToDevice));
        }
        *out = static_cast<const C*>(b.p);
        return MSX_OK;
    }
    ~msx_ctx_s() {
        for (auto& kv : tw32) kv.second.release();
        for (auto& kv : tw64) kv.second.release();
        drop_graphs();
        for (DevBuf* b : {&wave, &wave2, &trans, &aux, &tabs, &partial, &outbuf, &atoms, &cells, &vtmp, &curvebuf, &radtab,
                           &radcoef, &binbuf, &series})
            b->release();
        for (Lane& l : lanes) {
            l.wave.release();
            l.aux.release();
            l.tabs.release();
            if (l.done) hipEventDestroy(l.done);
            if (l.stream) hipStreamDestroy(l.stream);
        }
        if (lanes_go) hipEventDestroy(lanes_go);
        if (h_pin) hipHostFree(h_pin);
        if (ev0) hipEventDestroy(ev0);
        if (ev1) hipEventDestroy(ev1);
        if (ev_order) hipEventDestroy(ev_order);
        if (own) hipStreamDestroy(own);
    }
};

namespace {

int check_grid(int nx, int ny) {
    if (!supported_n(nx) || !supported_n(ny))
        return fail(MSX_ESIZE, "grid %d x %d: each axis must be one of 16,32,64,128,256,384,512,768,1024,2048", nx, ny);
    return MSX_OK;
}
int check_prec_where(int prec, int where) {
    if (prec != MSX_F32 && prec != MSX_F64) return fail(MSX_EINVAL, "prec must be MSX_F32 or MSX_F64");
    if (where != MSX_HOST && where != MSX_DEVICE) return fail(MSX_EINVAL, "where must be MSX_HOST or MSX_DEVICE");
    return MSX_OK;
}

// Per-call device tables for the slice loop: freqn k^2 tables and propagator
// factors per layer, packed in ctx->tabs.
template <typename C> struct SliceTables {
    const float* kx2 = nullptr;
    const float* ky2 = nullptr;
    const C* px = nullptr;  // [nlayer][nx]
    const C* py = nullptr;  // [nlayer][ny]
    float k2max = 0.f;
    // twiddles of the x and y transforms; null: the context's cached tables
    const C* twx = nullptr;
    const C* twy = nullptr;
};

// Host images of the slice tables: freqn k^2 tables (float, the reference's
// arithmetic) and the per-layer propagator factors, with their byte offsets
// in one 256-byte-aligned device block.
template <typename C> struct HostTables {
    std::vector<float> kx2, ky2;
    std::vector<C> px, py;
    size_t o_ky = 0, o_px = 0, o_py = 0, tot = 0;
};

inline size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

// Propagator factors along one axis of n points, cell length a, slice
// thickness cz, wavelength lam, tt = 2 tan(tilt): mulslice.cpp:494-513,
// t = pi*cz*(k^2*lambda - k*2tan(tilt)), P = (cos t, -sin t), evaluated in
// FP64 with k = index/a.
template <typename C> void prop_factors(int n, double a, double cz, double lam, double tt, C* out) {
    for (int i = 0; i < n; i++) {
        const double k = (double)fidx(i, n) / a;
        const double t = M_PI * cz * (k * k * lam - k * tt);
        out[i].x = std::cos(t);
        out[i].y = -std::sin(t);
    }
}

template <typename C> void host_tables(const msx_stack* st, bool with_prop, HostTables<C>* h) {
    const int nx = st->nx, ny = st->ny, nl = with_prop ? st->nlayer : 0;
    h->kx2 = k2_table(nx, st->ax);
    h->ky2 = k2_table(ny, st->by);
    h->px.assign((size_t)nl * nx, C{});
    h->py.assign((size_t)nl * ny, C{});
    const double lam = msx_wavelength(st->kev);
    const double tx = 2.0 * std::tan(st->ctilt_x), ty = 2.0 * std::tan(st->ctilt_y);
    for (int l = 0; l < nl; l++) {
        prop_factors<C>(nx, st->ax, st->cz[l], lam, tx, &h->px[(size_t)l * nx]);
        prop_factors<C>(ny, st->by, st->cz[l], lam, ty, &h->py[(size_t)l * ny]);
    }
    h->o_ky = al256(sizeof(float) * nx);
    h->o_px = h->o_ky + al256(sizeof(float) * ny);
    h->o_py = h->o_px + al256(sizeof(C) * h->px.size());
    h->tot = h->o_py + al256(sizeof(C) * h->py.size() + 1);
}

// Copies the host tables to the device block at `base` (at least h.tot
// bytes) on the context stream and waits for the copies.
template <typename C>
int upload_tables(msx_ctx ctx, const msx_stack* st, const HostTables<C>& h, char* base, SliceTables<C>* out) {
    const int nx = st->nx, ny = st->ny;
    HIPCHK(hipMemcpyAsync(base, h.kx2.data(), sizeof(float) * nx, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(base + h.o_ky, h.ky2.data(), sizeof(float) * ny, hipMemcpyHostToDevice, ctx->stream));
    if (!h.px.empty()) {
        HIPCHK(hipMemcpyAsync(base + h.o_px, h.px.data(), sizeof(C) * h.px.size(), hipMemcpyHostToDevice,
                              ctx->stream));
        HIPCHK(hipMemcpyAsync(base + h.o_py, h.py.data(), sizeof(C) * h.py.size(), hipMemcpyHostToDevice,
                              ctx->stream));
    }
    // the host vectors must outlive the async copies
    HIPCHK(hipStreamSynchronize(ctx->stream));
    out->kx2 = reinterpret_cast<const float*>(base);
    out->ky2 = reinterpret_cast<const float*>(base + h.o_ky);
    out->px = reinterpret_cast<const C*>(base + h.o_px);
    out->py = reinterpret_cast<const C*>(base + h.o_py);
    out->k2max = k2max_of(nx, ny, st->ax, st->by, st->bw);
    ctx->band = band_fraction(h.ky2.data(), ny, out->k2max);
    return MSX_OK;
}

template <typename C>
int build_tables(msx_ctx ctx, const msx_stack* st, bool with_prop, SliceTables<C>* out) {
    HostTables<C> h;
    host_tables<C>(st, with_prop, &h);
    CHK(ctx->tabs.ensure(h.tot));
    return upload_tables<C>(ctx, st, h, static_cast<char*>(ctx->tabs.p), out);
}

template <typename C> RowArgs<C> row_args(C* psi, int nx, int nb, const C* tw) {
    RowArgs<C> r;
    std::memset(&r, 0, sizeof r);
    r.psi = psi;
    r.nx = nx;
    r.nb = nb;
    r.tw = tw;
    r.scale = 1;
    return r;
}
template <typename C> ColArgs<C> col_args(C* psi, int ny, int nb, const C* tw) {
    ColArgs<C> c;
    std::memset(&c, 0, sizeof c);
    c.psi = psi;
    c.ny = ny;
    c.nb = nb;
    c.tw = tw;
    c.ops = kColStore;
    return c;
}

// Algorithmic HBM bytes of one sweep launch: every wave pixel read and
// written once (unless not stored), plus one shared transmission image.
template <typename C> double row_bytes(const RowArgs<C>& a, int n, double band) {
    using Re = typename real_of<C>::type;
    const double px = (double)a.nx * n;
    const double fin = (a.prune & 1) ? band : 1.0, fout = a.nostore ? 0.0 : ((a.prune & 2) ? band : 1.0);
    return px * a.nb * sizeof(C) * (fin + fout) + (a.trans ? px * sizeof(C) : 0.0) +
           (a.inten ? px * a.nb * sizeof(Re) * (a.iacc == 2 ? 2.0 : 1.0) : 0.0) +
           (a.trans_tab ? px * a.nb * sizeof(C) : 0.0) +
           (a.vreal ? px * a.nb * (0.5 * (double)sizeof(C) - (double)sizeof(C)) : 0.0);
}
template <typename C> double col_bytes(const ColArgs<C>& a, int n, double band) {
    const double px = (double)a.ny * n * a.nb * (a.prune ? band : 1.0);
    return px * (((a.ops & kColGen) ? 0.0 : sizeof(C)) + ((a.ops & kColStore) ? sizeof(C) : 0.0));
}
#define LAUNCH_ROW(n, mode, args)                                                                       \
    do {                                                                                                \
        if (ctx->keying) {                                                                              \
            ctx->key_add(MSX_PROF_ROW, (n), (mode), &(args), sizeof(args));                             \
            break;                                                                                      \
        }                                                                                               \
        CHK(ctx->timed(MSX_PROF_ROW, ctx->prof ? row_bytes<C>((args), (n), ctx->band) : 0.0,            \
                       [&] { return launch_row<C>((n), (mode), (args), ctx->stream); }));               \
    } while (0)
#define LAUNCH_COL(n, mode, args)                                                                       \
    do {                                                                                                \
        if (ctx->keying) {                                                                              \
            ctx->key_add(MSX_PROF_COL, (n), (mode), &(args), sizeof(args));                             \
            break;                                                                                      \
        }                                                                                               \
        CHK(ctx->timed(MSX_PROF_COL, ctx->prof ? col_bytes<C>((args), (n), ctx->band) : 0.0,            \
                       [&] { return launch_col<C>((n), (mode), (args), ctx->stream); }));               \
    } while (0)

// ---------------------------------------------------------------- fft2d
template <typename C> int fft2d_dev(msx_ctx ctx, C* d, int nb, int nx, int ny, int sign) {
    using Re = typename real_of<C>::type;
    const C *twx, *twy;
    CHK(ctx->twiddles<C>(nx, &twx));
    CHK(ctx->twiddles<C>(ny, &twy));
    RowArgs<C> r = row_args<C>(d, nx, nb, twy);
    ColArgs<C> c = col_args<C>(d, ny, nb, twx);
    if (sign > 0) {
        LAUNCH_ROW(ny, kRowFwd, r);
        LAUNCH_COL(nx, kColFwd, c);
    } else {
        r.scale = (Re)1 / ((Re)nx * (Re)ny);
        LAUNCH_ROW(ny, kRowInv0, r);
        LAUNCH_COL(nx, kColInv, c);
    }
    return MSX_OK;
}

// ---------------------------------------------------------------- transmission
template <typename C>
int transmission_dev(msx_ctx ctx, const typename real_of<C>::type* V, int nl, int nx, int ny, double ax, double by,
                     double kev, double bw, C* t, const typename real_of<C>::type* Vabs = nullptr) {
    using Re = typename real_of<C>::type;
    const C *twx, *twy;
    CHK(ctx->twiddles<C>(nx, &twx));
    CHK(ctx->twiddles<C>(ny, &twy));
    RowArgs<C> r = row_args<C>(t, nx, nl, twy);
    r.vreal = V;
    r.vabs = Vabs;
    r.vscale = (Re)(msx_sigma(kev) / 1000.0);
    if (!(bw > 0)) {
        LAUNCH_ROW(ny, kRowNone, r);
        return MSX_OK;
    }
    msx_stack st{};
    st.nx = nx;
    st.ny = ny;
    st.ax = ax;
    st.by = by;
    st.kev = kev;
    st.bw = bw;
    SliceTables<C> tb;
    CHK(build_tables<C>(ctx, &st, false, &tb));
    // Band pruning, exact as in the slice loop: the mask zeroes every y
    // column with ky2 >= k2max, so the forward row sweep does not store
    // those columns, the column sweep skips them, and the inverse row sweep
    // takes them as zeros without loading them (and stores the full image).
    r.ky2 = tb.ky2;
    r.k2max = tb.k2max;
    r.prune = 2;
    LAUNCH_ROW(ny, kRowFwd, r);
    ColArgs<C> c = col_args<C>(t, ny, nl, twx);
    c.kx2 = tb.kx2;
    c.ky2 = tb.ky2;
    c.k2max = tb.k2max;
    c.prune = 1;
    c.ops = kColMask | kColStore;
    LAUNCH_COL(nx, kColSlice, c);
    RowArgs<C> r2 = row_args<C>(t, nx, nl, twy);
    r2.scale = (Re)1 / ((Re)nx * (Re)ny);
    r2.ky2 = tb.ky2;
    r2.k2max = tb.k2max;
    r2.prune = 1;
    LAUNCH_ROW(ny, kRowInv, r2);
    return MSX_OK;
}

// ---------------------------------------------------------------- slice loop
struct DetPlan {
    int ndet = 0;
    float dlo[kMaxDet], dhi[kMaxDet];
    const int32_t* ck = nullptr;
    int nck = 0;
    double* out = nullptr;      // device [nck][ndet][nprobe]
    long long nprobe = 0, probe0 = 0;
};

// Device side of a thickness series (msx_multislice_series): what is
// recorded after each slice, read from the batch between the column sweep
// and the next row sweep.
template <typename C> struct SeriesPlan {
    const int32_t* at = nullptr;  // host: slice counts of the kept waves, ascending
    int nat = 0, iat = 0;         // count, next one to keep
    C* waves = nullptr;           // [nat][nb][nx][ny]
    const int4* bm = nullptr;     // [nbeams] (h, k, in band, 0)
    int nbeams = 0;
    C* beams = nullptr;           // [nslice][nb][nbeams]
    double* inten = nullptr;      // [nslice][nb], or nullptr
    double* partial = nullptr;    // [nb][nparts]
    int nparts = 1;
};

// The records of slice s (0-based): psi holds IFFT_x(P FFT_x FFT_y(t psi_s)),
// nx times the y transform of the exit wave of slice s.
template <typename C>
int series_records(msx_ctx ctx, const msx_stack* st, const SliceTables<C>& tb, C* psi, int nb, int s, const C* twy,
                   bool obl, SeriesPlan<C>* ser) {
    using Re = typename real_of<C>::type;
    const int nx = st->nx, ny = st->ny;
    const long long img = (long long)nx * ny;
    const double dnx = nx, dny = ny;
    if (ser->nbeams > 0)
        HIPCHK(launch_series_beams<C>(psi, nb, nx, ny, ser->bm, ser->nbeams, 1.0 / (dnx * dnx * dny),
                                      ser->beams + (long long)s * nb * ser->nbeams, ctx->stream));
    if (ser->inten) {
        // unwritten columns (ky2 >= k2max) are skipped unless the cell is
        // oblique (then every column was written)
        HIPCHK(launch_series_norm<C>(psi, nb, img, ny, obl ? nullptr : tb.ky2, tb.k2max, ser->nparts, ser->partial,
                                     ctx->stream));
        HIPCHK(launch_det_finish(ser->partial, 1, nb, ser->nparts, 1.0 / (dnx * dnx * dnx * dny * dny),
                                 ser->inten + (long long)s * nb, 0, 0, ctx->stream));
    }
    if (ser->iat < ser->nat && ser->at[ser->iat] == s + 1) {
        // the loop's final inverse row sweep, into the kept wave instead of psi
        RowArgs<C> r = row_args<C>(ser->waves + (long long)ser->iat * nb * img, nx, nb, twy);
        r.src = psi;
        r.src_bstride = img;
        r.scale = (Re)1 / ((Re)nx * (Re)ny);
        r.ky2 = tb.ky2;
        r.k2max = tb.k2max;
        r.prune = obl ? 0 : 1;
        LAUNCH_ROW(ny, kRowInv, r);
        ser->iat++;
    }
    return MSX_OK;
}

// Advances nb waves through the stack.  start: 0 = psi holds real-space waves
// (or src is a shared incident wave), 1 = psi holds k-space probes already
// through the column inverse (the first row sweep starts with IFFT_y).
// With det != nullptr the loop stops after the last checkpoint; otherwise
// it ends with the final inverse row sweep (exit waves in real space), or,
// with mixed_out, just before it: psi then holds IFFT_x(P FFT_x FFT_y(t psi))
// of the last slice, i.e. FFT_y of the exit wave times nx, in the columns
// ky2 < k2max (the others are not written).  ser: thickness-series records
// after every slice (no graph capture: never set on the keyed path).
template <typename C>
int slice_loop_body(msx_ctx ctx, const msx_stack* st, const SliceTables<C>& tb, const C* trans, C* psi, int nb,
                    const C* src, long long src_bstride, int start, DetPlan* det, bool mixed_out = false,
                    SeriesPlan<C>* ser = nullptr) {
    using Re = typename real_of<C>::type;
    const int nx = st->nx, ny = st->ny;
    const long long img = (long long)nx * ny;
    const C *twx = tb.twx, *twy = tb.twy;
    if (!twx) CHK(ctx->twiddles<C>(nx, &twx));
    if (!twy) CHK(ctx->twiddles<C>(ny, &twy));
    const Re inv = (Re)1 / ((Re)nx * (Re)ny);
    const msx_tuning* tu = ctx->tuning(nx, ny, sizeof(Re) == 8 ? MSX_F64 : MSX_F32);
    const int tile = tu ? tu->col_tile : 0;
    const int groups = col_groups<C>(nx, ny, tile);
    // oblique cell: the per-element FP64 propagator, no band pruning (the
    // columns beyond the band are not all-zero columns of the oblique disk)
    const bool obl = is_oblique(st);
    const double sg = obl ? std::sin(st->gam) : 1.0, cg = obl ? std::cos(st->gam) : 0.0;
    const double obl_k = std::min(nx / (2.0 * st->ax), ny / (2.0 * st->by)) * st->bw;
    int ick = 0;
    for (int s = 0; s < st->nslice; s++) {
        const int l = st->layer[s];
        // Band pruning (exact): after a column sweep every y column with
        // ky2 >= k2max is identically zero, so row sweeps neither load those
        // inputs (prune & 1, from the second slice on) nor store those outputs
        // (prune & 2: the next column sweep skips them).
        RowArgs<C> r = row_args<C>(psi, nx, nb, twy);
        r.trans = trans + (long long)l * img;
        r.ky2 = tb.ky2;
        r.k2max = tb.k2max;
        r.prune = obl ? 0 : ((s == 0) ? 2 : 3);
        if (tu) r.wpb = tu->row_wpb, r.plan = tu->row_plan;
        if (s == 0 && start == 0) {
            r.src = src;
            r.src_bstride = src_bstride;
            LAUNCH_ROW(ny, kRowFwd, r);
        } else {
            r.scale = inv;
            LAUNCH_ROW(ny, kRowSlice, r);
        }
        ColArgs<C> c = col_args<C>(psi, ny, nb, twx);
        c.prune = obl ? 0 : 1;
        c.kx2 = tb.kx2;
        c.ky2 = tb.ky2;
        c.k2max = tb.k2max;
        c.px = tb.px + (long long)l * nx;
        c.py = tb.py + (long long)l * ny;
        c.ops = kColProp | kColStore;
        if (tu) c.plan = tu->col_plan;
        if (obl) {
            // Propagator.cpp:65-72 with this slice's thickness
            const double lam = msx_wavelength(st->kev), dz = st->cz[l];
            c.ops |= kColObl;
            c.obl_asg = st->ax * sg;
            c.obl_bsg = st->by * sg;
            c.obl_2c = 2.0 * cg;
            c.obl_smax = obl_k * obl_k;
            c.obl_pl = M_PI * lam * dz;
            c.obl_tx = 2.0 * M_PI * dz * std::tan(st->ctilt_x);
            c.obl_ty = 2.0 * M_PI * dz * std::tan(st->ctilt_y);
        }
        if (tu) c.wpb = tu->col_wpb;
        c.tile = tile;
        bool is_ck = det && ick < det->nck && det->ck[ick] == s + 1;
        if (is_ck) {
            c.ops |= kColDetect;
            c.ndet = det->ndet;
            for (int d = 0; d < kMaxDet; d++) {
                c.dlo[d] = det->dlo[d];
                c.dhi[d] = det->dhi[d];
            }
            c.partial = static_cast<double*>(ctx->partial.p);
        }
        const bool last = det && is_ck && ick == det->nck - 1;
        if (last) {
            c.ops &= ~kColStore;  // nothing reads the wave after the last checkpoint
            LAUNCH_COL(nx, kColFwd, c);
        } else {
            LAUNCH_COL(nx, kColSlice, c);
        }
        if (is_ck) {
            double* dout = det->out + (long long)ick * det->ndet * det->nprobe;
            if (ctx->keying) {
                const long long fa[4] = {(long long)(intptr_t)c.partial, (long long)(intptr_t)dout, det->nprobe,
                                         det->probe0};
                ctx->key_add(MSX_PROF_DET, groups, det->ndet, fa, sizeof fa);
            } else {
                CHK(ctx->timed(MSX_PROF_DET, 8.0 * det->ndet * nb * (groups + 1), [&] {
                    return launch_det_finish(c.partial, det->ndet, nb, groups, 1.0 / ((double)nx * ny), dout,
                                             det->nprobe, det->probe0, ctx->stream);
                }));
            }
            ick++;
        }
        if (last) return MSX_OK;
        if (ser) CHK(series_records<C>(ctx, st, tb, psi, nb, s, twy, obl, ser));
    }
    if (!det && !mixed_out) {
        RowArgs<C> r = row_args<C>(psi, nx, nb, twy);
        r.scale = inv;
        r.ky2 = tb.ky2;
        r.k2max = tb.k2max;
        r.prune = obl ? 0 : 1;
        if (tu) r.plan = tu->row_plan;
        LAUNCH_ROW(ny, kRowInv, r);
    }
    return MSX_OK;
}

// The slice loop, replayed from a cached hipGraph when graphs are enabled
// (and profiling is off).  Kernel arguments are baked into a graph, so the
// key is the byte string of every launch's arguments: any change of pointer,
// size, table or detector set selects (or captures) another graph.  A capture
// that fails falls back to launching the loop directly.
template <typename C>
int slice_loop(msx_ctx ctx, const msx_stack* st, const SliceTables<C>& tb, const C* trans, C* psi, int nb,
               const C* src, long long src_bstride, int start, DetPlan* det, bool mixed_out = false) {
    if (!ctx->graphs || ctx->prof)
        return slice_loop_body<C>(ctx, st, tb, trans, psi, nb, src, src_bstride, start, det, mixed_out);
    ctx->gkey.clear();
    ctx->keying = true;
    const int rk = slice_loop_body<C>(ctx, st, tb, trans, psi, nb, src, src_bstride, start, det, mixed_out);
    ctx->keying = false;
    CHK(rk);
    auto it = ctx->gcache.find(ctx->gkey);
    if (it == ctx->gcache.end()) {
        if (ctx->gcache.size() >= 64) ctx->drop_graphs();
        hipGraph_t g = nullptr;
        if (hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            (void)hipGetLastError();  // e.g. the legacy default stream: launch directly
            return slice_loop_body<C>(ctx, st, tb, trans, psi, nb, src, src_bstride, start, det, mixed_out);
        }
        const int rc = slice_loop_body<C>(ctx, st, tb, trans, psi, nb, src, src_bstride, start, det, mixed_out);
        const hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
        hipGraphExec_t ex = nullptr;
        hipError_t ei = (rc == MSX_OK && ec == hipSuccess && g) ? hipGraphInstantiate(&ex, g, nullptr, nullptr, 0)
                                                                : hipErrorUnknown;
        if (g) hipGraphDestroy(g);
        if (ei != hipSuccess) {
            (void)hipGetLastError();
            return slice_loop_body<C>(ctx, st, tb, trans, psi, nb, src, src_bstride, start, det, mixed_out);
        }
        it = ctx->gcache.emplace(ctx->gkey, ex).first;
        ctx->graph_captures++;
    } else {
        ctx->graph_hits++;
    }
    HIPCHK(hipGraphLaunch(it->second, ctx->stream));
    return MSX_OK;
}

// ---------------------------------------------------------------- autotuner
// Times the slice loop on synthetic waves (zero data: the kernels do the same
// work whatever the values) for each candidate launch shape, with the
// per-launch event timing of the profiling mode, and keeps the fastest:
// For every STEM batch size (16 to 512 waves per launch, ten sizes): the waves per
// work-group of the row sweep and, for the column sweep, the tile width (the
// built-in one and its alternative) with its schedule: chunked with 1, 2, 4
// or 8 waves per work-group, or balanced with 1 or 2 work-groups per
// resident slot (each sweep on its own).  Among the batches within 10% of the
// best of those rates, the slice loop is then timed as it runs (launches back
// to back, events around the whole loop): the fastest batch is chosen, the
// smaller one within 1%, with its own shapes.
template <typename C> int autotune_impl(msx_ctx ctx, int nx, int ny, int max_batch, msx_tuning* res) {
    using Re = typename real_of<C>::type;
    const int prec = sizeof(Re) == 8 ? MSX_F64 : MSX_F32;
    const long long img = (long long)nx * ny;
    const long long cap = std::max(1LL, (2LL << 30) / (long long)(sizeof(C) * img));
    const int batches[] = {16, 32, 64, 96, 128, 160, 192, 256, 384, 512};
    constexpr int kNB = (int)(sizeof(batches) / sizeof(batches[0]));
    int bmax = (int)std::min<long long>(max_batch > 0 ? max_batch : 512, cap);
    const int nslice = 6;
    std::vector<double> cz(1, 2.0);
    std::vector<int32_t> lay(nslice, 0);
    msx_stack st{};
    st.nx = nx;
    st.ny = ny;
    st.ax = 0.1 * nx;
    st.by = 0.1 * ny;
    st.kev = 300.0;
    st.bw = 2.0 / 3.0;
    st.nlayer = 1;
    st.cz = cz.data();
    st.nslice = nslice;
    st.layer = lay.data();
    SliceTables<C> tb;
    CHK(build_tables<C>(ctx, &st, true, &tb));
    CHK(ctx->wave.ensure(sizeof(C) * img * bmax));
    CHK(ctx->trans.ensure(sizeof(C) * img));
    HIPCHK(hipMemsetAsync(ctx->wave.p, 0, sizeof(C) * img * bmax, ctx->stream));
    HIPCHK(hipMemsetAsync(ctx->trans.p, 0, sizeof(C) * img, ctx->stream));
    const bool prof0 = ctx->prof, graphs0 = ctx->graphs;
    if (!ctx->ev0) {
        HIPCHK(hipEventCreate(&ctx->ev0));
        HIPCHK(hipEventCreate(&ctx->ev1));
    }
    ctx->graphs = false;
    auto key = std::make_tuple(nx, ny, prec);
    const bool had = ctx->tune.count(key) > 0;
    const msx_tuning saved = had ? ctx->tune[key] : msx_tuning{};
    // times (ms per launch) of the row and column sweeps for one shape, on
    // the STEM loop's sequence (waves start in k-space, no final inverse row
    // sweep)
    auto trial = [&](int rw, int cw, int tile, int nb, double* row_ms, double* col_ms, int rp = 0,
                     int cp = 0) -> int {
        ctx->tune[key] = msx_tuning{rw, cw, 0, 0.0, tile, rp, cp};
        ctx->prof = true;
        for (int rep = 0; rep < 2; rep++) {  // the first pass warms up
            for (int k = 0; k < MSX_PROF_KINDS; k++) ctx->prof_n[k] = 0, ctx->prof_ms[k] = 0, ctx->prof_bytes[k] = 0;
            CHK(slice_loop_body<C>(ctx, &st, tb, static_cast<const C*>(ctx->trans.p), static_cast<C*>(ctx->wave.p),
                                   nb, nullptr, 0, 1, nullptr, true));
        }
        *row_ms = ctx->prof_ms[MSX_PROF_ROW] / std::max(1LL, ctx->prof_n[MSX_PROF_ROW]);
        *col_ms = ctx->prof_ms[MSX_PROF_COL] / std::max(1LL, ctx->prof_n[MSX_PROF_COL]);
        return MSX_OK;
    };
    int tdef, talt;
    col_tiles<C>(nx, &tdef, &talt);
    std::vector<int> tiles{tdef};
    if (talt != tdef) tiles.push_back(talt);
    int rc = MSX_OK;
    msx_tuning best{1, 1, 0, 0.0, 0, 0, 0};
    do {
        // per batch size: the fastest waves-per-work-group of the row sweep,
        // the fastest (tile, waves-per-work-group) of the column sweep, and
        // the resulting probe-slice rate
        double rate[kNB] = {0}, top = 0.0;
        int rw_of[kNB], cw_of[kNB], tile_of[kNB];
        for (int i = 0; i < kNB; i++) rw_of[i] = cw_of[i] = 1, tile_of[i] = tdef;
        for (int i = 0; i < kNB && rc == MSX_OK; i++) {
            if (batches[i] > bmax) break;
            double brow = 1e30, bcol = 1e30;
            for (int tile : tiles)
                for (int w : {1, 2, 4, 8, -1, -2}) {
                    // w < 0: the column sweep's balanced schedule (the row
                    // sweep is timed at positive w only)
                    double rm, cm;
                    if ((rc = trial(w > 0 ? w : 8, w, tile, batches[i], &rm, &cm)) != MSX_OK) break;
                    if (tile == tdef && w > 0 && rm < brow) brow = rm, rw_of[i] = w;
                    if (cm < bcol) bcol = cm, cw_of[i] = w, tile_of[i] = tile;
                }
            rate[i] = batches[i] / (brow + bcol);
            top = std::max(top, rate[i]);
        }
        if (rc != MSX_OK) break;
        // The per-launch rates above time each sweep alone (profiling mode
        // waits for every launch).  Among the batches within 10% of the best
        // such rate, time the slice loop as it runs (launches back to back,
        // events around the whole loop) and keep the fastest; the smaller
        // batch wins within 1%.
        auto loop_rate = [&](int i, double* out) -> int {
            ctx->tune[key] = msx_tuning{rw_of[i], cw_of[i], 0, 0.0, tile_of[i], 0, 0};
            ctx->prof = false;
            float ms = 0.f;
            for (int rep = 0; rep < 2; rep++) {  // the first pass warms up
                HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
                CHK(slice_loop_body<C>(ctx, &st, tb, static_cast<const C*>(ctx->trans.p),
                                       static_cast<C*>(ctx->wave.p), batches[i], nullptr, 0, 1, nullptr, true));
                HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
                HIPCHK(hipEventSynchronize(ctx->ev1));
                HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
            }
            *out = batches[i] * (double)nslice / std::max(1e-6, (double)ms);
            return MSX_OK;
        };
        double rbest = 0.0;
        for (int i = 0; i < kNB && batches[i] <= bmax && rc == MSX_OK; i++) {
            if (rate[i] < 0.9 * top) continue;
            double real = 0.0;
            if ((rc = loop_rate(i, &real)) != MSX_OK) break;
            if (real > 1.01 * rbest) {
                rbest = real;
                best = msx_tuning{rw_of[i], cw_of[i], batches[i], 1.0 / real, tile_of[i], 0, 0};
            }
        }
        if (rc != MSX_OK) break;
        // at that batch and with those shapes, each sweep's FFT plan (row and
        // column sweeps are timed separately, so one trial per variant times
        // both)
        // (row lines have length ny, column lines nx)
        int ralt = 0, calt = 0, unused = 0;
        plan_alts(ny, &ralt, &unused);
        plan_alts(nx, &unused, &calt);
        const int nplan = (ralt || calt) ? 2 : 1;
        if (nplan > 1) {
            double brow = 1e30, bcol = 1e30;
            int rbest = 0, cbest = 0;
            for (int v = 0; v < nplan && rc == MSX_OK; v++) {
                double rm, cm;
                if ((rc = trial(best.row_wpb, best.col_wpb, best.col_tile, best.batch, &rm, &cm, ralt ? v : 0,
                                calt ? v : 0)) != MSX_OK)
                    break;
                if (rm < brow) brow = rm, rbest = ralt ? v : 0;
                if (cm < bcol) bcol = cm, cbest = calt ? v : 0;
            }
            if (rc != MSX_OK) break;
            best.row_plan = rbest;
            best.col_plan = cbest;
            best.ms_per_slice = (brow + bcol) / best.batch;
        }
    } while (0);
    ctx->prof = prof0;
    ctx->graphs = graphs0;
    for (int k = 0; k < MSX_PROF_KINDS; k++) ctx->prof_n[k] = 0, ctx->prof_ms[k] = 0, ctx->prof_bytes[k] = 0;
    if (rc != MSX_OK) {
        if (had) ctx->tune[key] = saved; else ctx->tune.erase(key);
        return rc;
    }
    ctx->tune[key] = best;
    if (res) *res = best;
    return MSX_OK;
}

// ---------------------------------------------------------------- job batches
// Uneven HRTEM jobs (msx_multislice_batch).  Jobs on the same grid (size,
// cell and band limit: the tables of the band pruning) form one class,
// whatever their kV, tilt and slice thicknesses.  Within a class the waves
// are ordered by slice count (longest first), so the waves still running at
// slice s are a prefix, and each slice step is one row sweep (each wave's
// transmission read through a per-wave pointer table) and one column sweep.
// Where the running waves' propagators differ, the column sweep's persistent
// work-groups pull (wave, column group) units from a device-side counter and
// read each wave's propagator factors through a second pointer table (one
// factor array per distinct kV, tilt and thickness, the same FP64-evaluated
// values msx_multislice uses); where they agree it is the plain sweep with
// that array.  A job's final inverse row sweep runs right after its last
// slice.
struct BatchItem {
    int job;
    const msx_stack* st;
};

template <typename C>
int batch_class(msx_ctx ctx, const msx_job* jobs, std::vector<BatchItem>& g, std::vector<std::vector<char>>& keep,
                bool sync) {
    using Re = typename real_of<C>::type;
    std::stable_sort(g.begin(), g.end(),
                     [](const BatchItem& a, const BatchItem& b) { return a.st->nslice > b.st->nslice; });
    const msx_stack* ref = g[0].st;  // the longest
    const int nx = ref->nx, ny = ref->ny, nb = (int)g.size(), S = ref->nslice;
    const long long img = (long long)nx * ny;
    SliceTables<C> tb;
    CHK(build_tables<C>(ctx, ref, false, &tb));  // the class's k^2 tables and band limit
    std::vector<int> active(S + 1, 0);
    for (int s = 0; s <= S; s++)
        for (int i = 0; i < nb; i++) active[s] += g[i].st->nslice > s;
    // distinct propagators (kV, tilt, thickness) of the class's slices
    std::map<std::tuple<double, double, double, double>, int> pid;
    std::vector<int> pidx((size_t)S * nb, 0);
    for (int s = 0; s < S; s++)
        for (int i = 0; i < active[s]; i++) {
            const msx_stack* js = g[i].st;
            const auto key = std::make_tuple(js->kev, js->ctilt_x, js->ctilt_y, js->cz[js->layer[s]]);
            auto it = pid.find(key);
            if (it == pid.end()) it = pid.emplace(key, (int)pid.size()).first;
            pidx[(size_t)s * nb + i] = it->second;
        }
    const int U = (int)pid.size();
    // one device block: transmission pointers [S][nb], propagator pointers
    // [S][nb], the slices' work counters [S][kQueueSlots] (zero: one per XCD,
    // kQueueStride apart) and the factors [U][nx + ny]
    const size_t o_tt = 0, o_pt = al256(sizeof(void*) * S * nb), o_q = o_pt + al256(sizeof(void*) * S * nb),
                 o_pf = o_q + al256(sizeof(unsigned) * kQueueSlots * S),
                 tot = o_pf + sizeof(C) * (size_t)U * (nx + ny);
    CHK(ctx->aux.ensure(tot));
    char* dbase = static_cast<char*>(ctx->aux.p);
    keep.emplace_back(tot, 0);  // host image, kept by the caller until the stream has consumed it
    char* hbase = keep.back().data();
    const C** htt = reinterpret_cast<const C**>(hbase + o_tt);
    const C** hpt = reinterpret_cast<const C**>(hbase + o_pt);
    C* hpf = reinterpret_cast<C*>(hbase + o_pf);
    for (const auto& kv : pid) {
        const double kev = std::get<0>(kv.first), tx = std::get<1>(kv.first), ty = std::get<2>(kv.first);
        const double dz = std::get<3>(kv.first), lam = msx_wavelength(kev);
        C* f = hpf + (size_t)kv.second * (nx + ny);
        prop_factors<C>(nx, ref->ax, dz, lam, 2.0 * std::tan(tx), f);
        prop_factors<C>(ny, ref->by, dz, lam, 2.0 * std::tan(ty), f + nx);
    }
    const C* dpf = reinterpret_cast<const C*>(dbase + o_pf);
    for (int s = 0; s < S; s++)
        for (int i = 0; i < active[s]; i++) {
            const msx_stack* js = g[i].st;
            htt[(size_t)s * nb + i] = static_cast<const C*>(jobs[g[i].job].trans) + (long long)js->layer[s] * img;
            hpt[(size_t)s * nb + i] = dpf + (size_t)pidx[(size_t)s * nb + i] * (nx + ny);
        }
    HIPCHK(hipMemcpyAsync(dbase, hbase, tot, hipMemcpyHostToDevice, ctx->stream));
    const C* const* dtt = reinterpret_cast<const C* const*>(dbase + o_tt);
    const C* const* dpt = reinterpret_cast<const C* const*>(dbase + o_pt);
    unsigned* dq = reinterpret_cast<unsigned*>(dbase + o_q);
    // wave slots
    CHK(ctx->wave.ensure(sizeof(C) * img * nb));
    C* w = static_cast<C*>(ctx->wave.p);
    for (int i = 0; i < nb; i++) {
        const msx_job& j = jobs[g[i].job];
        HIPCHK(hipMemcpyAsync(w + i * img, j.psi0 ? j.psi0 : j.psi, sizeof(C) * img, hipMemcpyDeviceToDevice,
                              ctx->stream));
    }
    const C *twx, *twy;
    CHK(ctx->twiddles<C>(nx, &twx));
    CHK(ctx->twiddles<C>(ny, &twy));
    const Re inv = (Re)1 / ((Re)nx * (Re)ny);
    const msx_tuning* tu = ctx->tuning(nx, ny, sizeof(Re) == 8 ? MSX_F64 : MSX_F32);
    for (int s = 0; s < S; s++) {
        const int na = active[s];
        RowArgs<C> r = row_args<C>(w, nx, na, twy);
        r.trans_tab = dtt + (size_t)s * nb;
        r.ky2 = tb.ky2;
        r.k2max = tb.k2max;
        r.prune = (s == 0) ? 2 : 3;
        if (tu) r.wpb = tu->row_wpb, r.plan = tu->row_plan;
        if (s == 0) {
            LAUNCH_ROW(ny, kRowFwd, r);
        } else {
            r.scale = inv;
            LAUNCH_ROW(ny, kRowSlice, r);
        }
        ColArgs<C> c = col_args<C>(w, ny, na, twx);
        c.prune = 1;
        c.kx2 = tb.kx2;
        c.ky2 = tb.ky2;
        c.k2max = tb.k2max;
        c.tile = tu ? tu->col_tile : 0;
        if (tu) c.plan = tu->col_plan;
        // one propagator for every running wave (frozen-phonon configurations
        // of one specimen): the plain column sweep with that factor array;
        // otherwise the per-wave tables and the work queue
        bool uniform = true;
        for (int i = 1; i < na && uniform; i++) uniform = pidx[(size_t)s * nb + i] == pidx[(size_t)s * nb];
        if (uniform) {
            c.ops = kColProp | kColStore;
            c.px = dpf + (size_t)pidx[(size_t)s * nb] * (nx + ny);
            c.py = c.px + nx;
            if (tu) c.wpb = tu->col_wpb;
        } else {
            c.ops = kColProp | kColJobs | kColStore;
            c.prop_tab = dpt + (size_t)s * nb;
            c.queue = dq + (size_t)s * kQueueSlots;
            c.wpb = (tu && tu->col_wpb < 0) ? tu->col_wpb : -1;
        }
        LAUNCH_COL(nx, kColSlice, c);
        const int done = na - active[s + 1];  // jobs whose last slice this was
        if (done > 0) {
            RowArgs<C> f = row_args<C>(w + (long long)active[s + 1] * img, nx, done, twy);
            if (tu) f.plan = tu->row_plan;  // the single-job loop's final sweep
            f.scale = inv;
            f.ky2 = tb.ky2;
            f.k2max = tb.k2max;
            f.prune = 1;
            LAUNCH_ROW(ny, kRowInv, f);
        }
    }
    for (int i = 0; i < nb; i++)
        HIPCHK(hipMemcpyAsync(jobs[g[i].job].psi, w + i * img, sizeof(C) * img, hipMemcpyDeviceToDevice, ctx->stream));
    if (sync) HIPCHK(hipStreamSynchronize(ctx->stream));
    return MSX_OK;
}

int check_stack(const msx_stack* st);

template <typename C> int multislice_batch_impl(msx_ctx ctx, const msx_job* jobs, int njobs) {
    std::vector<bool> used(njobs, false);
    for (int i = 0; i < njobs; i++) CHK(check_stack(jobs[i].stack));
    auto same_grid = [](const msx_stack* a, const msx_stack* b) {
        return a->nx == b->nx && a->ny == b->ny && a->ax == b->ax && a->by == b->by && a->bw == b->bw;
    };
    std::vector<std::vector<BatchItem>> groups;
    for (int i = 0; i < njobs; i++) {
        if (used[i]) continue;
        std::vector<BatchItem> g{{i, jobs[i].stack}};
        used[i] = true;
        for (int k = i + 1; k < njobs; k++) {
            if (used[k] || !same_grid(jobs[i].stack, jobs[k].stack)) continue;
            g.push_back({k, jobs[k].stack});
            used[k] = true;
        }
        groups.push_back(std::move(g));
    }
    std::vector<std::vector<char>> keep;
    // Groups are independent: up to kLanes of them run at once, each on its
    // own stream (groups of small grids alone leave most of the GPU idle).
    // Profiling times launches one at a time, so it keeps one stream.
    constexpr int kLanes = 4;
    const int nl = ctx->prof ? 1 : std::min<int>(kLanes, (int)groups.size());
    if (nl <= 1) {
        for (auto& g : groups) CHK(batch_class<C>(ctx, jobs, g, keep, true));
        return MSX_OK;
    }
    CHK(ctx->ensure_lanes(nl));
    // twiddle tables are shared: create them before any lane runs
    for (auto& g : groups) {
        const C* tw;
        CHK(ctx->twiddles<C>(g[0].st->nx, &tw));
        CHK(ctx->twiddles<C>(g[0].st->ny, &tw));
    }
    HIPCHK(hipEventRecord(ctx->lanes_go, ctx->stream));
    for (int l = 0; l < nl; l++) HIPCHK(hipStreamWaitEvent(ctx->lanes[l].stream, ctx->lanes_go, 0));
    // longest work first, each class to the lane with the least work so far
    // (work: pixel-slices of the class's jobs)
    std::vector<std::pair<double, size_t>> order;
    for (size_t gi = 0; gi < groups.size(); gi++) {
        double w = 0.0;
        for (const auto& it : groups[gi]) w += (double)it.st->nslice * it.st->nx * it.st->ny;
        order.emplace_back(w, gi);
    }
    std::stable_sort(order.begin(), order.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
    std::vector<double> load(nl, 0.0);
    int rc = MSX_OK;
    for (size_t k = 0; k < order.size() && rc == MSX_OK; k++) {
        const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        load[l] += order[k].first;
        ctx->swap_lane(l);
        rc = batch_class<C>(ctx, jobs, groups[order[k].second], keep, false);
        ctx->swap_lane(l);
    }
    // the context stream waits for every lane; the call returns when all are done
    for (int l = 0; l < nl; l++) {
        HIPCHK(hipEventRecord(ctx->lanes[l].done, ctx->lanes[l].stream));
        HIPCHK(hipStreamWaitEvent(ctx->stream, ctx->lanes[l].done, 0));
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return rc;
}

int check_stack(const msx_stack* st) {
    if (!st) return fail(MSX_EINVAL, "stack is NULL");
    if (!(st->gam == 0.0 || (st->gam > 0.0 && st->gam < M_PI)))
        return fail(MSX_EINVAL, "gam = %g: the cell angle must be 0 (orthogonal) or in (0, pi)", st->gam);
    CHK(check_grid(st->nx, st->ny));
    if (!(st->ax > 0) || !(st->by > 0) || !(st->kev > 0)) return fail(MSX_EINVAL, "ax, by, kev must be > 0");
    if (st->nlayer <= 0 || !st->cz) return fail(MSX_EINVAL, "nlayer must be > 0 with cz set");
    if (st->nslice <= 0 || !st->layer) return fail(MSX_EINVAL, "nslice must be > 0 with layer set");
    for (int s = 0; s < st->nslice; s++)
        if (st->layer[s] < 0 || st->layer[s] >= st->nlayer)
            return fail(MSX_EINVAL, "layer[%d] = %d out of range [0,%d)", s, st->layer[s], st->nlayer);
    return MSX_OK;
}

template <typename C>
int multislice_impl(msx_ctx ctx, const msx_stack* st, const void* trans, const void* psi0, void* psi, int nwave,
                    int where) {
    const long long img = (long long)st->nx * st->ny;
    const size_t tbytes = sizeof(C) * img * st->nlayer, wbytes = sizeof(C) * img * nwave;
    const C* t_dev = static_cast<const C*>(trans);
    C* w_dev = static_cast<C*>(psi);
    const C* s_dev = static_cast<const C*>(psi0);
    if (where == MSX_HOST) {
        CHK(ctx->trans.ensure(tbytes));
        CHK(ctx->wave.ensure(wbytes));
        HIPCHK(hipMemcpyAsync(ctx->trans.p, trans, tbytes, hipMemcpyHostToDevice, ctx->stream));
        t_dev = static_cast<const C*>(ctx->trans.p);
        w_dev = static_cast<C*>(ctx->wave.p);
        if (psi0) {
            CHK(ctx->aux.ensure(sizeof(C) * img));
            HIPCHK(hipMemcpyAsync(ctx->aux.p, psi0, sizeof(C) * img, hipMemcpyHostToDevice, ctx->stream));
            s_dev = static_cast<const C*>(ctx->aux.p);
        } else {
            HIPCHK(hipMemcpyAsync(w_dev, psi, wbytes, hipMemcpyHostToDevice, ctx->stream));
        }
    }
    SliceTables<C> tb;
    CHK(build_tables<C>(ctx, st, true, &tb));
    CHK(slice_loop<C>(ctx, st, tb, t_dev, w_dev, nwave, s_dev, 0, 0, nullptr));
    if (where == MSX_HOST) HIPCHK(hipMemcpyAsync(psi, w_dev, wbytes, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return MSX_OK;
}

// msx_multislice with the records of a thickness series.  One device block
// (ctx->series) holds the beam table, the intensity partial sums and, for
// host outputs, the device copies of the kept waves, beams and intensities.
template <typename C>
int multislice_series_impl(msx_ctx ctx, const msx_stack* st, const void* trans, const void* psi0, void* psi,
                           int nwave, int where, const msx_series* se) {
    const int nx = st->nx, ny = st->ny;
    const long long img = (long long)nx * ny;
    const size_t tbytes = sizeof(C) * img * st->nlayer, wbytes = sizeof(C) * img * nwave;
    const bool obl = is_oblique(st);
    // beams: mulslice.cpp:693-700 (wrap a negative index once, then clamp),
    // flagged out of band with the column sweep's float test; an oblique
    // cell's band is the FP64 metric disk: every beam is transformed there
    const std::vector<float> kx2 = k2_table(nx, st->ax), ky2 = k2_table(ny, st->by);
    const float k2max = k2max_of(nx, ny, st->ax, st->by, st->bw);
    std::vector<int4> bm((size_t)std::max(se->nbeams, 1));
    for (int i = 0; i < se->nbeams; i++) {
        int h = se->hk[2 * i], k = se->hk[2 * i + 1];
        if (h < 0) h += nx;
        if (k < 0) k += ny;
        h = std::min(std::max(h, 0), nx - 1);
        k = std::min(std::max(k, 0), ny - 1);
        bm[i] = make_int4(h, k, (obl || kx2[h] + ky2[k] < k2max) ? 1 : 0, 0);
    }
    const int nparts =
        se->intensity ? (int)std::max(1LL, std::min<long long>(std::min(1024 / nwave, 1024), img / 4096)) : 1;
    const bool host = where == MSX_HOST;
    const size_t o_bm = 0, o_part = al256(sizeof(int4) * bm.size()),
                 o_wav = o_part + al256(sizeof(double) * nwave * nparts),
                 o_bea = o_wav + (host ? al256(sizeof(C) * img * nwave * se->nat) : 0),
                 o_int = o_bea + (host ? al256(sizeof(C) * (size_t)st->nslice * nwave * se->nbeams) : 0),
                 tot = o_int + (host && se->intensity ? al256(sizeof(double) * st->nslice * nwave) : 0);
    CHK(ctx->series.ensure(tot));
    char* base = static_cast<char*>(ctx->series.p);
    HIPCHK(hipMemcpyAsync(base + o_bm, bm.data(), sizeof(int4) * bm.size(), hipMemcpyHostToDevice, ctx->stream));
    SeriesPlan<C> sp;
    sp.at = se->at;
    sp.nat = se->nat;
    sp.bm = reinterpret_cast<const int4*>(base + o_bm);
    sp.nbeams = se->nbeams;
    sp.partial = reinterpret_cast<double*>(base + o_part);
    sp.nparts = nparts;
    sp.waves = static_cast<C*>(host ? (void*)(base + o_wav) : se->waves);
    sp.beams = static_cast<C*>(host ? (void*)(base + o_bea) : se->beams);
    sp.inten = se->intensity ? (host ? reinterpret_cast<double*>(base + o_int) : se->intensity) : nullptr;

    const C* t_dev = static_cast<const C*>(trans);
    C* w_dev = static_cast<C*>(psi);
    const C* s_dev = static_cast<const C*>(psi0);
    if (host) {
        CHK(ctx->trans.ensure(tbytes));
        CHK(ctx->wave.ensure(wbytes));
        HIPCHK(hipMemcpyAsync(ctx->trans.p, trans, tbytes, hipMemcpyHostToDevice, ctx->stream));
        t_dev = static_cast<const C*>(ctx->trans.p);
        w_dev = static_cast<C*>(ctx->wave.p);
        if (psi0) {
            CHK(ctx->aux.ensure(sizeof(C) * img));
            HIPCHK(hipMemcpyAsync(ctx->aux.p, psi0, sizeof(C) * img, hipMemcpyHostToDevice, ctx->stream));
            s_dev = static_cast<const C*>(ctx->aux.p);
        } else {
            HIPCHK(hipMemcpyAsync(w_dev, psi, wbytes, hipMemcpyHostToDevice, ctx->stream));
        }
    }
    SliceTables<C> tb;
    CHK(build_tables<C>(ctx, st, true, &tb));  // waits for the copies above (bm outlives them)
    CHK(slice_loop_body<C>(ctx, st, tb, t_dev, w_dev, nwave, s_dev, 0, 0, nullptr, false, &sp));
    if (host) {
        HIPCHK(hipMemcpyAsync(psi, w_dev, wbytes, hipMemcpyDeviceToHost, ctx->stream));
        if (se->nat > 0)
            HIPCHK(hipMemcpyAsync(se->waves, sp.waves, sizeof(C) * img * nwave * se->nat, hipMemcpyDeviceToHost,
                                  ctx->stream));
        if (se->nbeams > 0)
            HIPCHK(hipMemcpyAsync(se->beams, sp.beams, sizeof(C) * (size_t)st->nslice * nwave * se->nbeams,
                                  hipMemcpyDeviceToHost, ctx->stream));
        if (se->intensity)
            HIPCHK(hipMemcpyAsync(se->intensity, sp.inten, sizeof(double) * st->nslice * nwave,
                                  hipMemcpyDeviceToHost, ctx->stream));
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return MSX_OK;
}

// msx_multislice on device buffers with the slice tables and the twiddles
// in the GPU's `scratch` array: no allocation, no graph capture (a graph
// instantiation may allocate); the same kernels and tables as msx_multislice.
template <typename C>
int multislice_noalloc_impl(msx_ctx ctx, const msx_stack* st, const void* trans, const void* psi0, void* psi,
                            int nwave) {
    HostTables<C> h;
    host_tables<C>(st, true, &h);
    const size_t o_twx = h.tot, o_twy = o_twx + al256(sizeof(C) * st->nx),
                 need = o_twy + al256(sizeof(C) * st->ny);
    if (need > (size_t)MSX_NOALLOC_SCRATCH_BYTES)
        return fail(MSX_ESIZE,
                    "msx_multislice_noalloc: the tables of %d layers on a %d x %d grid take %zu bytes, more than "
                    "the %zu-byte scratch array (MSX_NOALLOC_SCRATCH_BYTES)",
                    st->nlayer, st->nx, st->ny, need, (size_t)MSX_NOALLOC_SCRATCH_BYTES);
    if (!ctx->scratch) return fail(MSX_ESTATE, "msx_multislice_noalloc: no scratch array on device %d", ctx->device);
    char* base = static_cast<char*>(ctx->scratch);
    const std::vector<C> twx = make_twiddles<C>(st->nx), twy = make_twiddles<C>(st->ny);
    HIPCHK(hipMemcpyAsync(base + o_twx, twx.data(), sizeof(C) * st->nx, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(base + o_twy, twy.data(), sizeof(C) * st->ny, hipMemcpyHostToDevice, ctx->stream));
    SliceTables<C> tb;
    CHK(upload_tables<C>(ctx, st, h, base, &tb));  // waits for every copy above
    tb.twx = reinterpret_cast<const C*>(base + o_twx);
    tb.twy = reinterpret_cast<const C*>(base + o_twy);
    CHK(slice_loop_body<C>(ctx, st, tb, static_cast<const C*>(trans), static_cast<C*>(psi), nwave,
                           static_cast<const C*>(psi0), 0, 0, nullptr));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return MSX_OK;
}

template <typename C>
int stem_impl(msx_ctx ctx, const msx_stack* st, const void* trans, const msx_probe* pr, const double* pos,
              int nprobe, const msx_detector* det, int ndet, const int32_t* ck, int nck, double weight,
              int accumulate, double* out, int batch, int where) {
    using Re = typename real_of<C>::type;
    const int nx = st->nx, ny = st->ny;
    const long long img = (long long)nx * ny;
    const msx_tuning* tu = ctx->tuning(nx, ny, sizeof(Re) == 8 ? MSX_F64 : MSX_F32);
    if (batch <= 0 && tu && tu->batch > 0) batch = tu->batch;
    if (batch <= 0) {
        // default: enough waves to fill the GPU, bounded by ~2 GiB of waves
        long long b = (2LL << 30) / (long long)(sizeof(C) * img);
        batch = (int)std::max(1LL, std::min<long long>(b, 256));
    }
    batch = std::min(batch, nprobe);
    const C* t_dev = static_cast<const C*>(trans);
    if (where == MSX_HOST) {
        const size_t tbytes = sizeof(C) * img * st->nlayer;
        CHK(ctx->trans.ensure(tbytes));
        HIPCHK(hipMemcpyAsync(ctx->trans.p, trans, tbytes, hipMemcpyHostToDevice, ctx->stream));
        t_dev = static_cast<const C*>(ctx->trans.p);
    }
    CHK(ctx->wave.ensure(sizeof(C) * img * batch));
    C* w_dev = static_cast<C*>(ctx->wave.p);
    const int groups = col_groups<C>(nx, ny, tu ? tu->col_tile : 0);
    CHK(ctx->partial.ensure(sizeof(double) * (size_t)ndet * batch * groups));
    // device: results [nck][ndet][nprobe], the current batch's results
    // [nck][ndet][batch] (one address for every batch, so the slice-loop graph
    // of every full batch is the same) and all probe positions
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t obytes = sizeof(double) * (size_t)nck * ndet * nprobe;
    const size_t sbytes = sizeof(double) * (size_t)nck * ndet * batch;
    CHK(ctx->outbuf.ensure(al(obytes) + al(sbytes) + sizeof(double2) * (size_t)nprobe));
    double* o_dev = static_cast<double*>(ctx->outbuf.p);
    double* s_dev = reinterpret_cast<double*>(static_cast<char*>(ctx->outbuf.p) + al(obytes));
    double2* pos_dev = reinterpret_cast<double2*>(static_cast<char*>(ctx->outbuf.p) + al(obytes) + al(sbytes));

    SliceTables<C> tb;
    CHK(build_tables<C>(ctx, st, true, &tb));
    const double lam = msx_wavelength(st->kev);
    const float kap_hi = (float)((pr->alpha_max / lam) * (pr->alpha_max / lam));
    const float kap_lo = (float)((pr->alpha_min / lam) * (pr->alpha_min / lam));
    // aperture pixel count with the same float test as the kernel
    std::vector<float> kx2 = k2_table(nx, st->ax), ky2 = k2_table(ny, st->by);
    long long nap = 0;
    for (int ix = 0; ix < nx; ix++)
        for (int iy = 0; iy < ny; iy++) {
            float k2 = kx2[ix] + ky2[iy];
            if (k2 <= kap_hi && k2 >= kap_lo) nap++;
        }
    if (nap == 0) return fail(MSX_EINVAL, "probe aperture contains no pixel");
    const C *twx;
    CHK(ctx->twiddles<C>(nx, &twx));

    DetPlan dp;
    dp.ndet = ndet;
    for (int d = 0; d < kMaxDet; d++) {
        dp.dlo[d] = d < ndet ? (float)(det[d].k_inner * det[d].k_inner) : 1.f;
        dp.dhi[d] = d < ndet ? (float)(det[d].k_outer * det[d].k_outer) : 0.f;
    }
    dp.ck = ck;
    dp.nck = nck;
    dp.out = s_dev;
    dp.nprobe = batch;
    dp.probe0 = 0;
    // all positions in one copy: batches follow each other without a host wait
    std::vector<double2> hpos(nprobe);
    for (int p = 0; p < nprobe; p++) {
        hpos[p].x = pos[2 * p];
        hpos[p].y = pos[2 * p + 1];
    }
    HIPCHK(hipMemcpyAsync(pos_dev, hpos.data(), sizeof(double2) * nprobe, hipMemcpyHostToDevice, ctx->stream));
    for (int p0 = 0; p0 < nprobe; p0 += batch) {
        const int nb = std::min(batch, nprobe - p0);
        ColArgs<C> g = col_args<C>(w_dev, ny, nb, twx);
        g.ops = kColGen | kColStore;
        g.kx2 = tb.kx2;
        g.ky2 = tb.ky2;
        g.pos = pos_dev + p0;
        g.ax = st->ax;
        g.by = st->by;
        g.chi1 = M_PI * lam;
        g.chi2 = 0.5 * pr->cs * lam * lam;
        g.df = pr->df;
        g.kap_lo = kap_lo;
        g.kap_hi = kap_hi;
        g.amp = (Re)std::sqrt((double)img / (double)nap);
        if (tu) g.plan = tu->col_plan;
        LAUNCH_COL(nx, kColGenInv, g);
        CHK(slice_loop<C>(ctx, st, tb, t_dev, w_dev, nb, nullptr, 0, 1, &dp));
        // this batch's [nck][ndet][nb] block into columns p0.. of the results
        HIPCHK(hipMemcpy2DAsync(o_dev + p0, sizeof(double) * nprobe, s_dev, sizeof(double) * batch,
                                sizeof(double) * nb, (size_t)nck * ndet, hipMemcpyDeviceToDevice, ctx->stream));
    }
    std::vector<double> res((size_t)nck * ndet * nprobe);
    HIPCHK(hipMemcpyAsync(res.data(), o_dev, obytes, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (size_t i = 0; i < res.size(); i++) out[i] = (accumulate ? out[i] : 0.0) + weight * res[i];
    return MSX_OK;
}

template <typename C>
int ctf_impl(msx_ctx ctx, const void* psi, int nwave, int nx, int ny, double ax, double by, double kev, double cs,
             double df, double aobj, void* psi_img, void* inten, int where) {
    using Re = typename real_of<C>::type;
    const long long img = (long long)nx * ny;
    const size_t wbytes = sizeof(C) * img * nwave, ibytes = sizeof(Re) * img * nwave;
    CHK(ctx->wave2.ensure(wbytes));
    C* w = static_cast<C*>(ctx->wave2.p);
    const C* in = static_cast<const C*>(psi);
    Re* i_dev = static_cast<Re*>(inten);
    if (where == MSX_HOST) {
        HIPCHK(hipMemcpyAsync(w, psi, wbytes, hipMemcpyHostToDevice, ctx->stream));
        in = w;
        if (inten) {
            CHK(ctx->vtmp.ensure(ibytes));
            i_dev = static_cast<Re*>(ctx->vtmp.p);
        }
    }
    const C *twx, *twy;
    CHK(ctx->twiddles<C>(nx, &twx));
    CHK(ctx->twiddles<C>(ny, &twy));
    msx_stack st{};
    st.nx = nx;
    st.ny = ny;
    st.ax = ax;
    st.by = by;
    st.kev = kev;
    st.bw = 1.0;
    SliceTables<C> tb;
    CHK(build_tables<C>(ctx, &st, false, &tb));
    const double lam = msx_wavelength(kev);
    RowArgs<C> r = row_args<C>(w, nx, nwave, twy);
    r.src = in;
    r.src_bstride = img;
    LAUNCH_ROW(ny, kRowFwd, r);
    ColArgs<C> c = col_args<C>(w, ny, nwave, twx);
    c.ops = kColCtf | kColStore;
    c.kx2 = tb.kx2;
    c.ky2 = tb.ky2;
    c.ax = ax;
    c.by = by;
    c.chi1 = M_PI * lam;
    c.chi2 = 0.5 * cs * lam * lam;
    c.df = df;
    c.kap_hi = (float)((aobj / lam) * (aobj / lam));
    LAUNCH_COL(nx, kColSlice, c);
    RowArgs<C> r2 = row_args<C>(w, nx, nwave, twy);
    r2.scale = (Re)1 / ((Re)nx * (Re)ny);
    r2.inten = inten ? i_dev : nullptr;
    if (psi_img && where == MSX_DEVICE) {
        r2.src = w;
        r2.src_bstride = img;
        r2.psi = static_cast<C*>(psi_img);
    }
    LAUNCH_ROW(ny, kRowInv, r2);
    if (where == MSX_HOST) {
        if (psi_img) HIPCHK(hipMemcpyAsync(psi_img, w, wbytes, hipMemcpyDeviceToHost, ctx->stream));
        if (inten) HIPCHK(hipMemcpyAsync(inten, i_dev, ibytes, hipMemcpyDeviceToHost, ctx->stream));
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return MSX_OK;
}

// Illumination angles (iqx, iqy) of the partial-coherence integral in the
// float arithmetic of mulslice.cpp:556-583 (loop iqy outer, iqx inner).
std::vector<int2> illum_angles(const msx_stack* st, double acmin, double acmax) {
    const float wavlen = (float)msx_wavelength((double)(float)st->kev);
    const float rx = 1.0F / (float)st->ax, ry = 1.0F / (float)st->by;
    const float acmaxf = (float)acmax, acminf = (float)acmin;
    const int nacx = (int)((acmaxf / (wavlen * rx)) + 1.5F);
    const int nacy = (int)((acmaxf / (wavlen * ry)) + 1.5F);
    float q2max = acmaxf / wavlen;
    q2max = q2max * q2max;
    float q2min = acminf / wavlen;
    q2min = q2min * q2min;
    std::vector<int2> out;
    for (int iqy = -nacy; iqy <= nacy; iqy++) {
        const float qy = iqy * ry;
        const float qy2 = qy * qy;
        for (int iqx = -nacx; iqx <= nacx; iqx++) {
            const float qx = iqx * rx;
            const float q2 = qx * qx + qy2;
            if ((q2 <= q2max) && (q2 >= q2min)) out.push_back(make_int2(iqx, iqy));
        }
    }
    return out;
}

template <typename C>
int partial_impl(msx_ctx ctx, const msx_stack* st, const void* trans, const msx_lens* lens, double acmin,
                 double acmax, double sigmaf, double dfdelt, int batch, double* image, int32_t* nangles,
                 int32_t* nsteps, int where) {
    using Re = typename real_of<C>::type;
    const int nx = st->nx, ny = st->ny;
    const long long img = (long long)nx * ny;
    const std::vector<int2> iq = illum_angles(st, acmin, acmax);
    if (iq.empty()) return fail(MSX_EINVAL, "no illumination angle of the grid lies in [acmin, acmax]");
    // defocus steps (mulslice.cpp:556,617-645): ndf from the float expression,
    // weights exp(-0.5 ((df - df0)/sigmaf)^2)
    std::vector<double> off, wt;
    if (sigmaf > 0) {
        const int ndf = (int)((2.5F * (float)sigmaf) / (float)dfdelt);
        for (int i = -ndf; i <= ndf; i++) {
            const double x = (i * dfdelt) / sigmaf;
            off.push_back(i * dfdelt);
            wt.push_back(std::exp(-0.5 * x * x));
        }
    } else {
        off.push_back(0.0);
        wt.push_back(1.0);
    }
    double wsum = 0.0;
    for (double w : wt) wsum += w;
    const int nang = (int)iq.size();
    if (nangles) *nangles = nang;
    if (nsteps) *nsteps = (int)off.size();

    const msx_tuning* tu = ctx->tuning(nx, ny, sizeof(Re) == 8 ? MSX_F64 : MSX_F32);
    if (batch <= 0 && tu && tu->batch > 0) batch = tu->batch;
    if (batch <= 0) {
        // two wave buffers and one intensity buffer: about 2 GiB in all
        long long b = (2LL << 30) / (long long)((2 * sizeof(C) + sizeof(Re)) * img);
        batch = (int)std::max(1LL, std::min<long long>(b, 256));
    }
    batch = std::min(batch, nang);
    const C* t_dev = static_cast<const C*>(trans);
    if (where == MSX_HOST) {
        const size_t tbytes = sizeof(C) * img * st->nlayer;
        CHK(ctx->trans.ensure(tbytes));
        HIPCHK(hipMemcpyAsync(ctx->trans.p, trans, tbytes, hipMemcpyHostToDevice, ctx->stream));
        t_dev = static_cast<const C*>(ctx->trans.p);
    }
    CHK(ctx->wave.ensure(sizeof(C) * img * batch));
    CHK(ctx->wave2.ensure(sizeof(C) * img * batch));
    CHK(ctx->vtmp.ensure(sizeof(Re) * img * batch));
    C* w = static_cast<C*>(ctx->wave.p);
    C* w2 = static_cast<C*>(ctx->wave2.p);
    Re* ib = static_cast<Re*>(ctx->vtmp.p);
    const size_t ibytes = sizeof(double) * img, iq_off = (ibytes + 255) & ~(size_t)255;
    CHK(ctx->outbuf.ensure(iq_off + sizeof(int2) * iq.size()));
    double* pix = where == MSX_DEVICE ? image : static_cast<double*>(ctx->outbuf.p);
    int2* iq_dev = reinterpret_cast<int2*>(static_cast<char*>(ctx->outbuf.p) + iq_off);
    HIPCHK(hipMemcpyAsync(iq_dev, iq.data(), sizeof(int2) * iq.size(), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemsetAsync(pix, 0, ibytes, ctx->stream));

    SliceTables<C> tb;
    CHK(build_tables<C>(ctx, st, true, &tb));
    const C *twx, *twy;
    CHK(ctx->twiddles<C>(nx, &twx));
    CHK(ctx->twiddles<C>(ny, &twy));
    const double lam = msx_wavelength(st->kev);
    const float kap_hi = (float)((lens->aobj / lam) * (lens->aobj / lam));
    // columns with ky2 > kap_hi are zero after the lens and those with
    // ky2 >= k2max were not written by the slice loop: neither is touched
    const float k2p = std::min(tb.k2max, std::nextafter(kap_hi, INFINITY));
    // fractions of live columns (for the profiler's byte counts)
    const double band_slice = ctx->band, band_lens = band_fraction(k2_table(ny, st->by).data(), ny, k2p);
    const double scale = 1.0 / ((double)nang * wsum);
    for (int a0 = 0; a0 < nang; a0 += batch) {
        const int nb = std::min(batch, nang - a0);
        HIPCHK(launch_tilt_waves<C>(w, nb, nx, ny, iq_dev + a0, ctx->stream));
        // exit waves as nx * FFT_y(psi) (the slice loop without its last row sweep)
        ctx->band = band_slice;
        CHK(slice_loop<C>(ctx, st, tb, t_dev, w, nb, nullptr, 0, 0, nullptr, true));
        ctx->band = band_lens;
        for (size_t d = 0; d < off.size(); d++) {
            // FFT_x, lens at this defocus, IFFT_x: the spectrum Psi = FFT_x FFT_y psi
            // is this data over nx
            ColArgs<C> c = col_args<C>(w, ny, nb, twx);
            c.out = w2;
            c.ops = kColCtf | kColStore;
            c.prune = 1;
            c.kx2 = tb.kx2;
            c.ky2 = tb.ky2;
            c.k2max = k2p;
            c.ax = st->ax;
            c.by = st->by;
            c.chi1 = M_PI * lam;
            c.chi2 = 0.5 * lens->cs * lam * lam;
            c.df = lens->df + off[d];
            c.kap_hi = kap_hi;
            if (lens->dfa2 != 0.0 || lens->dfa3 != 0.0) {
                c.astig = 1;
                c.a2 = lens->dfa2;
                c.phi2 = lens->dfa2phi;
                c.a3 = 2.0 * lens->dfa3 * lam / 3.0;
                c.phi3 = lens->dfa3phi;
            }
            LAUNCH_COL(nx, kColSlice, c);
            // IFFT_y with the 1/(nx ny) of ifft and the 1/nx above; weighted
            // intensity summed over the steps per wave
            RowArgs<C> r = row_args<C>(w2, nx, nb, twy);
            r.scale = (Re)1 / ((Re)nx * (Re)nx * (Re)ny);
            r.ky2 = tb.ky2;
            r.k2max = k2p;
            r.prune = 1;
            r.inten = ib;
            r.iacc = d == 0 ? 1 : 2;
            r.iw = (Re)wt[d];
            r.nostore = 1;
            LAUNCH_ROW(ny, kRowInv, r);
        }
        HIPCHK(launch_wave_sum<Re>(ib, nb, img, scale, pix, ctx->stream));
    }
    if (where == MSX_HOST) HIPCHK(hipMemcpyAsync(image, pix, ibytes, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return MSX_OK;
}

// host binning of atoms for the splat kernel
// fn(lo, hi) over [0, n) split across up to 16 host threads (one below 32768
// items per thread); the ranges are disjoint, so fn may write per-item output
template <typename F> void parallel_ranges(int64_t n, F&& fn) {
    const int64_t hw = std::max(1u, std::thread::hardware_concurrency());
    const int T = (int)std::max<int64_t>(1, std::min<int64_t>(std::min<int64_t>(hw, 16), n / 32768));
    if (T == 1) {
        fn(int64_t(0), n);
        return;
    }
    std::vector<std::thread> th;
    th.reserve(T);
    for (int t = 0; t < T; t++) th.emplace_back([&fn, n, t, T] { fn(n * t / T, n * (t + 1) / T); });
    for (auto& x : th) x.join();
}

// lateral bins of the splat: cells of at least the cutoff radius, at most
// 1024 per axis
BinGeom splat_geom(double a, double b, double c, int nslices, double radius) {
    BinGeom g;
    g.a = a;
    g.b = b;
    g.dz = c / nslices;
    g.nslices = nslices;
    g.nbx = std::max(1, std::min(1024, (int)std::floor(a / radius)));
    g.nby = std::max(1, std::min(1024, (int)std::floor(b / radius)));
    return g;
}

int bad_atom(int64_t i, int Z) {
    if (Z < 1 || Z > 103) return fail(MSX_EINVAL, "atom %lld: Z = %d outside 1..103", (long long)i, Z);
    return fail(MSX_EINVAL, "atom %lld: coordinate not finite", (long long)i);
}

// radial tables (FP32 output), splat of the binned atoms in ctx->atoms /
// ctx->cells, copy to the host when asked
int splat_binned(msx_ctx ctx, RadArgs& ra, const BinGeom& g, int nx, int ny, double radius, int table,
                 int formula, void* V, int prec, int where) {
    const size_t vbytes = (prec == MSX_F64 ? sizeof(double) : sizeof(float)) * (size_t)g.nslices * nx * ny;
    void* v_dev = V;
    if (where == MSX_HOST) {
        CHK(ctx->vtmp.ensure(vbytes));
        v_dev = ctx->vtmp.p;
    }
    SplatArgs sa;
    sa.cell_start = static_cast<const int*>(ctx->cells.p);
    sa.atoms = static_cast<const AtomRec*>(ctx->atoms.p);
    sa.v64 = prec == MSX_F64 ? static_cast<double*>(v_dev) : nullptr;
    sa.v32 = prec == MSX_F32 ? static_cast<float*>(v_dev) : nullptr;
    sa.a = g.a;
    sa.b = g.b;
    sa.radius = radius;
    sa.nx = nx;
    sa.ny = ny;
    sa.nslices = g.nslices;
    sa.nbx = g.nbx;
    sa.nby = g.nby;
    sa.table = table;
    sa.formula = formula;
    sa.rtab = nullptr;
    sa.rcoef = nullptr;
    sa.rt_noct = 0;
    if (prec == MSX_F32 && ra.nsp > 0) {
        // FP32 output: V_Z(r^2) from radial tables (below FP32 resolution of
        // the exact sum, which the FP64 output keeps); octaves from 2^kRadOmin
        // up to the one holding radius^2
        ra.noct = std::max(1, (int)std::ceil(std::log2(radius * radius)) - kRadOmin + 1);
        ra.table = table;
        ra.formula = formula;
        CHK(ctx->radtab.ensure(sizeof(double) * (size_t)ra.nsp * ((size_t)ra.noct * (kRadPts + 3) + 1)));
        ra.tab = static_cast<double*>(ctx->radtab.p);
        CHK(ctx->radcoef.ensure(2 * sizeof(double2) * (size_t)ra.nsp * ra.noct * kRadPts));
        ra.coef = static_cast<double2*>(ctx->radcoef.p);
        HIPCHK(launch_radial_tables(ra, ctx->stream));
        sa.rtab = ra.tab;
        sa.rcoef = ra.coef;
        sa.rt_noct = ra.noct;
    }
    CHK(ctx->timed(MSX_PROF_SPLAT, (double)vbytes, [&] { return launch_splat(sa, ctx->stream); }));
    if (where == MSX_HOST) HIPCHK(hipMemcpyAsync(V, v_dev, vbytes, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return MSX_OK;
}

int potential_impl(msx_ctx ctx, const msx_atom* atoms, int64_t natoms, double a, double b, double c, int nx, int ny,
                   int nslices, double radius, int table, int formula, void* V, int prec, int where) {
    const BinGeom g = splat_geom(a, b, c, nslices, radius);
    const long long ncell = (long long)nslices * g.nbx * g.nby;
    std::vector<int>& cnt = ctx->h_cnt;
    std::vector<long long>& cell_of = ctx->h_cell;
    std::vector<AtomRec>& rec = ctx->h_rec;
    cnt.assign(ncell + 1, 0);
    cell_of.resize(natoms);
    rec.resize(natoms);
    // species present (radial tables: slots in ascending Z)
    int slot_of[104];
    std::fill(slot_of, slot_of + 104, -1);
    RadArgs ra;
    std::memset(&ra, 0, sizeof ra);
    bool present[104] = {};
    for (int64_t i = 0; i < natoms; i++) {
        const msx_atom& at = atoms[i];
        if (at.Z < 1 || at.Z > 103 || !std::isfinite(at.x) || !std::isfinite(at.y) || !std::isfinite(at.z))
            return bad_atom(i, at.Z);
        present[at.Z] = true;
    }
    for (int Z = 1; Z <= 103; Z++)
        if (present[Z]) {
            slot_of[Z] = ra.nsp;
            ra.zs[ra.nsp++] = Z;
        }
    // slice and lateral bin of every atom, in parallel over atom ranges
    parallel_ranges(natoms, [&](int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; i++) {
            const msx_atom& at = atoms[i];
            double x, y;
            if (!bin_atom(g, at.x, at.y, at.z, &cell_of[i], &x, &y)) {
                cell_of[i] = -1;
                continue;
            }
            rec[i] = AtomRec{x, y, at.occ, (int16_t)at.Z, (int16_t)slot_of[at.Z]};
        }
    });
    // counting sort by cell, atoms of a cell in input order
    for (int64_t i = 0; i < natoms; i++)
        if (cell_of[i] >= 0) cnt[cell_of[i] + 1]++;
    for (long long k = 0; k < ncell; k++) cnt[k + 1] += cnt[k];
    std::vector<int>& fill = ctx->h_fill;
    fill.assign(cnt.begin(), cnt.end() - 1);
    const size_t nsorted = (size_t)std::max<long long>(1, cnt[ncell]);
    // an earlier call that failed after its upload was queued may still be
    // reading the pinned records
    HIPCHK(hipStreamSynchronize(ctx->stream));
    CHK(ctx->ensure_pinned(nsorted));
    AtomRec* sorted = ctx->h_pin;
    for (int64_t i = 0; i < natoms; i++)
        if (cell_of[i] >= 0) sorted[fill[cell_of[i]]++] = rec[i];
    CHK(ctx->atoms.ensure(sizeof(AtomRec) * nsorted));
    CHK(ctx->cells.ensure(sizeof(int) * cnt.size()));
    // the pinned records are reused by the next call only after this call's
    // final stream synchronisation
    HIPCHK(hipMemcpyAsync(ctx->atoms.p, sorted, sizeof(AtomRec) * nsorted, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->cells.p, cnt.data(), sizeof(int) * cnt.size(), hipMemcpyHostToDevice, ctx->stream));
    return splat_binned(ctx, ra, g, nx, ny, radius, table, formula, V, prec, where);
}

// device atoms: natoms records inside one device allocation of this context's
// GPU (checked before any kernel reads them)
int check_device_atoms(msx_ctx ctx, const msx_atom* atoms, int64_t natoms) {
    if (natoms == 0) return MSX_OK;
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, atoms) != hipSuccess || attr.type != hipMemoryTypeDevice ||
        attr.device != ctx->device) {
        (void)hipGetLastError();
        return fail(MSX_EINVAL, "atoms: not device memory of GPU %d", ctx->device);
    }
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, const_cast<msx_atom*>(atoms)) != hipSuccess ||
        reinterpret_cast<const char*>(atoms) + natoms * sizeof(msx_atom) > static_cast<const char*>(base) + size) {
        (void)hipGetLastError();
        return fail(MSX_EINVAL, "atoms: %lld records run past the end of their allocation", (long long)natoms);
    }
    return MSX_OK;
}

// The same potential from atoms in device memory: binned on the GPU (stable
// radix sort by cell), so the records and cell starts equal the host
// binning's and the potential is identical.
int potential_dev_impl(msx_ctx ctx, const msx_atom* atoms, int64_t natoms, double a, double b, double c, int nx,
                       int ny, int nslices, double radius, int table, int formula, void* V, int prec, int where) {
    const BinGeom g = splat_geom(a, b, c, nslices, radius);
    const long long ncell = (long long)nslices * g.nbx * g.nby;
    if (natoms >= INT_MAX || ncell >= INT_MAX)
        return fail(MSX_EINVAL, "device atoms: %lld atoms, %lld cells (at most %d each)", (long long)natoms, ncell,
                    INT_MAX - 1);
    CHK(check_device_atoms(ctx, atoms, natoms));
    const int n = (int)natoms;
    size_t tb = 0;
    HIPCHK(bin_sort_temp_bytes(n, (int)ncell, &tb));
    auto up = [](size_t x) { return (x + 255) & ~size_t(255); };
    const size_t arr = up(sizeof(int) * std::max(n, 1));
    CHK(ctx->binbuf.ensure(4 * arr + up(sizeof(int) * 104) + up(std::max<size_t>(tb, 1))));
    char* base = static_cast<char*>(ctx->binbuf.p);
    DevBinArgs da;
    da.atoms = reinterpret_cast<const AtomIO*>(atoms);
    da.n = n;
    da.g = g;
    da.ncell = (int)ncell;
    da.key = reinterpret_cast<unsigned*>(base);
    da.val = reinterpret_cast<int*>(base + arr);
    da.key2 = reinterpret_cast<unsigned*>(base + 2 * arr);
    da.val2 = reinterpret_cast<int*>(base + 3 * arr);
    da.flags = reinterpret_cast<int*>(base + 4 * arr);
    da.temp = base + 4 * arr + up(sizeof(int) * 104);
    da.temp_bytes = tb;
    HIPCHK(launch_bin_keys(da, ctx->stream));
    int flags[104];
    HIPCHK(hipMemcpyAsync(flags, da.flags, sizeof flags, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (flags[0] != INT_MAX) {
        msx_atom at;
        HIPCHK(hipMemcpyAsync(&at, atoms + flags[0], sizeof at, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        return bad_atom(flags[0], at.Z);
    }
    RadArgs ra;
    std::memset(&ra, 0, sizeof ra);
    SlotMap slots;
    for (int Z = 0; Z < 104; Z++) slots.s[Z] = -1;
    for (int Z = 1; Z <= 103; Z++)
        if (flags[Z]) {
            slots.s[Z] = (int16_t)ra.nsp;
            ra.zs[ra.nsp++] = Z;
        }
    CHK(ctx->atoms.ensure(sizeof(AtomRec) * std::max(n, 1)));
    CHK(ctx->cells.ensure(sizeof(int) * (ncell + 1)));
    HIPCHK(launch_bin_sort(da, static_cast<int*>(ctx->cells.p), static_cast<AtomRec*>(ctx->atoms.p), slots,
                           ctx->stream));
    return splat_binned(ctx, ra, g, nx, ny, radius, table, formula, V, prec, where);
}

// 2 pi a0 e (V A^2): projected potential in V A of the Fourier sum of f_e/area
// (Kirkland's vzatom constant al = 4 pi^2 a0 e = 300.8242834, slicelib.cpp:2666-2699)
constexpr double kVA2 = 300.8242834 / (2.0 * 3.14159265358979323846);

// tabulated scattering curves (Muslib *_curve_* potentials): per-atom type
// index into curves[ntypes][npoints], spacing dds
struct CurveSpec {
    const int32_t* type;
    const double* curves;
    int ntypes, npoints;
    double dds;
};
int check_curves(const CurveSpec* cs, int64_t natoms) {
    if (!cs->type || !cs->curves || cs->ntypes < 1 || cs->npoints < 2 || !(cs->dds > 0) ||
        (long long)cs->ntypes * cs->npoints > (1LL << 26))
        return fail(MSX_EINVAL, "curves: type, table (ntypes >= 1, npoints >= 2) and dds > 0 required");
    for (int64_t i = 0; i < natoms; i++)
        if (cs->type[i] < 0 || cs->type[i] >= cs->ntypes)
            return fail(MSX_EINVAL, "atom %lld: type %d outside 0..%d", (long long)i, cs->type[i], cs->ntypes - 1);
    return MSX_OK;
}
int upload_curves(msx_ctx ctx, const CurveSpec* cs, const double** dev) {
    const size_t bytes = sizeof(double) * (size_t)cs->ntypes * cs->npoints;
    CHK(ctx->curvebuf.ensure(bytes));
    HIPCHK(hipMemcpyAsync(ctx->curvebuf.p, cs->curves, bytes, hipMemcpyHostToDevice, ctx->stream));
    *dev = static_cast<const double*>(ctx->curvebuf.p);
    return MSX_OK;
}

int potential_rec_impl(msx_ctx ctx, const msx_atom* atoms, int64_t natoms, const double* dw, const double* ads,
                       double a, double b, double c, double gam, int nmult, int mmult, int nx, int ny, int nslices,
                       void* V, void* Vabs, int prec, int where, const CurveSpec* cs = nullptr) {
    const double dz = c / nslices;
    const double sg = std::sin(gam), cg = std::cos(gam);
    struct Item {
        long long kz;
        int Z;
        double dw, ads;
        RecAtom at;
    };
    std::vector<Item> items;
    items.reserve(natoms);
    for (int64_t i = 0; i < natoms; i++) {
        const msx_atom& at = atoms[i];
        if (!cs && (at.Z < 1 || at.Z > 92))
            return fail(MSX_EINVAL, "atom %lld: Z = %d outside 1..92 (Doyle-Turner table)", (long long)i, at.Z);
        long long kz = (long long)std::floor(at.z / dz);
        if (kz >= 0 && kz < nslices && !(kz * dz <= at.z && at.z < (kz + 1) * dz)) kz += (at.z < kz * dz) ? -1 : 1;
        if (kz < 0 || kz >= nslices) continue;
        // fractional coordinates: a along x, b at angle gam from a in the xy plane
        const double yf = at.y / (b * sg);
        const double xf = (at.x - at.y * cg / sg) / a;
        items.push_back(Item{kz, cs ? cs->type[i] : at.Z, (dw && !cs) ? dw[i] : 0.0, ads ? ads[i] : 0.0,
                             RecAtom{xf - std::floor(xf), yf - std::floor(yf), (double)at.occ}});
    }
    std::sort(items.begin(), items.end(), [](const Item& p, const Item& q) {
        if (p.kz != q.kz) return p.kz < q.kz;
        if (p.Z != q.Z) return p.Z < q.Z;
        if (p.dw != q.dw) return p.dw < q.dw;
        return p.ads < q.ads;
    });
    std::vector<RecAtom> rat(std::max<size_t>(1, items.size()));
    std::vector<RecSeg> segs;
    std::vector<int> slice_seg(nslices + 1, 0);
    for (size_t i = 0; i < items.size(); i++) {
        rat[i] = items[i].at;
        const bool same = !segs.empty() && i > 0 && items[i - 1].kz == items[i].kz && items[i - 1].Z == items[i].Z &&
                          items[i - 1].dw == items[i].dw && items[i - 1].ads == items[i].ads;
        if (same) {
            segs.back().end = (int)i + 1;
        } else {
            segs.push_back(RecSeg{(int)i, (int)i + 1, items[i].Z, 0, items[i].dw, items[i].ads});
            slice_seg[items[i].kz + 1]++;
        }
    }
    for (int k = 0; k < nslices; k++) slice_seg[k + 1] += slice_seg[k];
    if (segs.empty()) segs.push_back(RecSeg{0, 0, 1, 0, 0.0, 0.0});
    CHK(ctx->atoms.ensure(sizeof(RecAtom) * rat.size() + sizeof(RecSeg) * segs.size()));
    CHK(ctx->cells.ensure(sizeof(int) * slice_seg.size()));
    char* abase = static_cast<char*>(ctx->atoms.p);
    RecSeg* d_segs = reinterpret_cast<RecSeg*>(abase + sizeof(RecAtom) * rat.size());
    HIPCHK(hipMemcpyAsync(abase, rat.data(), sizeof(RecAtom) * rat.size(), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(d_segs, segs.data(), sizeof(RecSeg) * segs.size(), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->cells.p, slice_seg.data(), sizeof(int) * slice_seg.size(), hipMemcpyHostToDevice,
                          ctx->stream));
    // slices in chunks of at most 1 GiB of complex128 coefficients
    const size_t img = (size_t)nx * ny;
    const int chunk = (int)std::max<size_t>(1, std::min<size_t>(nslices, ((size_t)1 << 30) / (img * 16)));
    CHK(ctx->wave2.ensure(img * 16 * chunk));
    const size_t rb = prec == MSX_F64 ? 8 : 4;
    if (where == MSX_HOST) CHK(ctx->vtmp.ensure(rb * img * chunk * (Vabs ? 2 : 1)));
    RecArgs ra;
    ra.atoms = reinterpret_cast<const RecAtom*>(abase);
    ra.segs = d_segs;
    ra.slice_seg = static_cast<const int*>(ctx->cells.p);
    ra.out = static_cast<double2*>(ctx->wave2.p);
    ra.A = 1.0 / (a * sg);
    ra.B = 1.0 / (b * sg);
    ra.cg = cg;
    ra.inv_area = 1.0 / (a * b * sg);
    ra.nx = nx;
    ra.ny = ny;
    ra.nmult = nmult;
    ra.mmult = mmult;
    ra.curves = nullptr;
    ra.npoints = 0;
    ra.dds = 1.0;
    if (cs) {
        CHK(upload_curves(ctx, cs, &ra.curves));
        ra.npoints = cs->npoints;
        ra.dds = cs->dds;
    }
    for (int k0 = 0; k0 < nslices; k0 += chunk) {
        const int nk = std::min(chunk, nslices - k0);
        ra.kz0 = k0;
        CHK(ctx->timed(MSX_PROF_SPLAT, (double)img * nk * 16, [&] { return launch_rec_sf(ra, nk, ctx->stream); }));
        // V(r) = sum_g V(g) exp(+2 pi i g.r): the unscaled +i transform
        CHK(fft2d_dev<double2>(ctx, ra.out, nk, nx, ny, +1));
        char* vd = where == MSX_HOST ? static_cast<char*>(ctx->vtmp.p) : static_cast<char*>(V) + rb * img * k0;
        char* va = nullptr;
        if (Vabs) va = where == MSX_HOST ? vd + rb * img * nk : static_cast<char*>(Vabs) + rb * img * k0;
        HIPCHK(launch_split_potential(ra.out, (long long)img * nk, kVA2, vd, va, prec == MSX_F64, ctx->stream));
        if (where == MSX_HOST) {
            HIPCHK(hipMemcpyAsync(static_cast<char*>(V) + rb * img * k0, vd, rb * img * nk, hipMemcpyDeviceToHost,
                                  ctx->stream));
            if (Vabs)
                HIPCHK(hipMemcpyAsync(static_cast<char*>(Vabs) + rb * img * k0, va, rb * img * nk,
                                      hipMemcpyDeviceToHost, ctx->stream));
            HIPCHK(hipStreamSynchronize(ctx->stream));
        }
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return MSX_OK;
}

int potential_rs_impl(msx_ctx ctx, const msx_atom* atoms, int64_t natoms, const double* dw, const double* ads,
                      double a, double b, double c, double gam, double atrad, int periodic, int nx, int ny,
                      int nslices, void* V, void* Vabs, int prec, int where, const CurveSpec* cs = nullptr) {
    const double dz = c / nslices;
    const double sg = std::sin(gam), cg = std::cos(gam);
    const double drax = atrad * sg / a, dray = atrad * sg / b;
    const int ntx = (nx + 15) / 16, nty = (ny + 15) / 16;
    const long long ntiles = (long long)nslices * ntx * nty;
    std::vector<RsAtom> rat;
    std::vector<int> akz;
    rat.reserve(natoms);
    for (int64_t i = 0; i < natoms; i++) {
        const msx_atom& at = atoms[i];
        if (!cs && (at.Z < 1 || at.Z > 92))
            return fail(MSX_EINVAL, "atom %lld: Z = %d outside 1..92 (Doyle-Turner table)", (long long)i, at.Z);
        long long kz = (long long)std::floor(at.z / dz);
        if (kz >= 0 && kz < nslices && !(kz * dz <= at.z && at.z < (kz + 1) * dz)) kz += (at.z < kz * dz) ? -1 : 1;
        if (kz < 0 || kz >= nslices) continue;
        const double yf0 = at.y / (b * sg), xf0 = (at.x - at.y * cg / sg) / a;
        const double xf = xf0 - std::floor(xf0), yf = yf0 - std::floor(yf0);
        // Pot_RS.cpp:131-134: C conversions truncate toward zero
        const long ibb = (long)((xf - drax) * nx), iee = (long)((xf + drax) * nx);
        const long jbb = (long)((yf - dray) * ny), jee = (long)((yf + dray) * ny);
        if (ibb < -(1L << 30) || iee > (1L << 30) || jbb < -(1L << 30) || jee > (1L << 30))
            return fail(MSX_EINVAL, "atrad too large for the grid");
        rat.push_back(RsAtom{xf, yf, (double)at.occ, (dw && !cs) ? dw[i] : 0.0, ads ? ads[i] : 0.0,
                             cs ? cs->type[i] : at.Z, (int)ibb, (int)iee, (int)jbb, (int)jee, 0});
        akz.push_back((int)kz);
    }
    // (atom, image) entries per (slice, tile): count, then fill
    std::vector<int> start(ntiles + 1, 0);
    auto boxes = [&](size_t k, auto&& emit) {
        const RsAtom& r = rat[k];
        for (int l = 0; l < (periodic ? 9 : 1); l++) {
            const int sx = (l == 3 || l == 5 || l == 8) ? 1 : ((l == 4 || l == 6 || l == 7) ? -1 : 0);
            const int sy = (l == 1 || l == 5 || l == 7) ? 1 : ((l == 2 || l == 6 || l == 8) ? -1 : 0);
            if ((sx == 1 && !(r.ibb < 0)) || (sx == -1 && !(r.iee >= nx)) || (sy == 1 && !(r.jbb < 0)) ||
                (sy == -1 && !(r.jee >= ny)))
                continue;
            int ib = std::max(r.ibb, 0), ie = std::min(r.iee, nx), jb = std::max(r.jbb, 0), je = std::min(r.jee, ny);
            if (sx == 1) { ib = r.ibb + nx; ie = nx; }
            if (sx == -1) { ib = 0; ie = r.iee - nx + 1; }
            if (sy == 1) { jb = r.jbb + ny; je = ny; }
            if (sy == -1) { jb = 0; je = r.jee - ny + 1; }
            ib = std::max(ib, 0);
            jb = std::max(jb, 0);
            ie = std::min(ie, nx);
            je = std::min(je, ny);
            if (ib >= ie || jb >= je) continue;
            for (int tx = ib / 16; tx <= (ie - 1) / 16; tx++)
                for (int ty = jb / 16; ty <= (je - 1) / 16; ty++)
                    emit(((long long)akz[k] * ntx + tx) * nty + ty, l);
        }
    };
    for (size_t k = 0; k < rat.size(); k++) boxes(k, [&](long long t, int) { start[t + 1]++; });
    for (long long t = 0; t < ntiles; t++) start[t + 1] += start[t];
    if (start[ntiles] > (1 << 28)) return fail(MSX_EINVAL, "%d box-tile entries: atrad too large", start[ntiles]);
    std::vector<int2> ent(std::max(1, start[ntiles]));
    {
        std::vector<int> fill(start.begin(), start.end() - 1);
        for (size_t k = 0; k < rat.size(); k++) boxes(k, [&](long long t, int l) { ent[fill[t]++] = make_int2((int)k, l); });
    }
    if (rat.empty()) rat.push_back(RsAtom{});
    const size_t abytes = sizeof(RsAtom) * rat.size(), ebytes = sizeof(int2) * ent.size();
    CHK(ctx->atoms.ensure(abytes + ebytes));
    CHK(ctx->cells.ensure(sizeof(int) * start.size()));
    char* base = static_cast<char*>(ctx->atoms.p);
    HIPCHK(hipMemcpyAsync(base, rat.data(), abytes, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(base + abytes, ent.data(), ebytes, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->cells.p, start.data(), sizeof(int) * start.size(), hipMemcpyHostToDevice, ctx->stream));
    const size_t rb = prec == MSX_F64 ? 8 : 4;
    const size_t vbytes = rb * (size_t)nslices * nx * ny;
    void* vd = V;
    void* vad = Vabs;
    if (where == MSX_HOST) {
        CHK(ctx->vtmp.ensure(vbytes * (Vabs ? 2 : 1)));
        vd = ctx->vtmp.p;
        vad = Vabs ? static_cast<char*>(ctx->vtmp.p) + vbytes : nullptr;
    }
    RsArgs ra;
    std::memset(&ra, 0, sizeof ra);
    ra.atoms = reinterpret_cast<const RsAtom*>(base);
    ra.entries = reinterpret_cast<const int2*>(base + abytes);
    ra.tile_start = static_cast<const int*>(ctx->cells.p);
    if (prec == MSX_F64) {
        ra.v64 = static_cast<double*>(vd);
        ra.a64 = static_cast<double*>(vad);
    } else {
        ra.v32 = static_cast<float*>(vd);
        ra.a32 = static_cast<float*>(vad);
    }
    ra.a = a;
    ra.b = b;
    ra.cg = cg;
    ra.dxa = a / nx;
    ra.dyb = b / ny;
    ra.scale = kVA2;
    ra.nx = nx;
    ra.ny = ny;
    ra.ntx = ntx;
    ra.nty = nty;
    ra.nslices = nslices;
    ra.dds = 1.0;
    if (cs) {
        CHK(upload_curves(ctx, cs, &ra.curves));
        ra.npoints = cs->npoints;
        ra.dds = cs->dds;
    }
    CHK(ctx->timed(MSX_PROF_SPLAT, (double)vbytes, [&] { return launch_rs_dt(ra, ctx->stream); }));
    if (where == MSX_HOST) {
        HIPCHK(hipMemcpyAsync(V, vd, vbytes, hipMemcpyDeviceToHost, ctx->stream));
        if (Vabs) HIPCHK(hipMemcpyAsync(Vabs, vad, vbytes, hipMemcpyDeviceToHost, ctx->stream));
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return MSX_OK;
}

}  // namespace

int msx_ctx_device(const msx_ctx_s* ctx) { return ctx->device; }
hipStream_t msx_ctx_stream(const msx_ctx_s* ctx) { return ctx->stream; }

// ============================================================== C ABI
extern "C" {

MSX_API int msx_version(void) { return 1; }
MSX_API const char* msx_last_error(void) { return g_err.c_str(); }

MSX_API int msx_create(int device, msx_ctx* out) {
    if (!out) return fail(MSX_EINVAL, "out is NULL");
    *out = nullptr;
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(MSX_EINVAL, "device %d not present (%d devices)", device, n);
    HIPCHK(hipSetDevice(device));
    msx_ctx c = new msx_ctx_s();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(MSX_EHIP, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    c->stream = c->own;
    if (scratch_address(device, &c->scratch) != MSX_OK) {
        delete c;
        return MSX_EHIP;
    }
    e = upload_fparams(c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        delete c;
        return fail(MSX_EHIP, "fparams upload: %s", hipGetErrorString(e));
    }
    *out = c;
    return MSX_OK;
}

MSX_API int msx_destroy(msx_ctx ctx) {
    if (!ctx) return MSX_OK;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    delete ctx;
    return MSX_OK;
}

MSX_API int msx_profile(msx_ctx ctx, int enable) {
    if (!ctx) return fail(MSX_ESTATE, "ctx is NULL");
    HIPCHK(hipSetDevice(ctx->device));
    if (enable && !ctx->ev0) {
        HIPCHK(hipEventCreate(&ctx->ev0));
        HIPCHK(hipEventCreate(&ctx->ev1));
    }
    ctx->prof = enable != 0;
    for (int k = 0; k < MSX_PROF_KINDS; k++) {
        ctx->prof_n[k] = 0;
        ctx->prof_ms[k] = 0;
        ctx->prof_bytes[k] = 0;
    }
    return MSX_OK;
}

MSX_API int msx_profile_read(msx_ctx ctx, int kind, int64_t* launches, double* total_ms, double* total_bytes) {
    if (!ctx) return fail(MSX_ESTATE, "ctx is NULL");
    if (kind < 0 || kind >= MSX_PROF_KINDS) return fail(MSX_EINVAL, "kind %d out of range", kind);
    if (launches) *launches = ctx->prof_n[kind];
    if (total_ms) *total_ms = ctx->prof_ms[kind];
    if (total_bytes) *total_bytes = ctx->prof_bytes[kind];
    return MSX_OK;
}

MSX_API int msx_copy_probe(msx_ctx ctx, int mode, const void* src, void* dst, int nimg, int nx, int ny,
                           double* ms) {
    if (!ctx) return fail(MSX_ESTATE, "ctx is NULL");
    if (!src || !dst || !ms || nimg < 1 || nx < 1 || ny < 1) return fail(MSX_EINVAL, "copy probe: bad arguments");
    if (mode != 0 && !(nx == 512 && (mode == 8 || mode == 16 || mode == 1008) && ny % (mode % 1000) == 0))
        return fail(MSX_EINVAL, "copy probe: mode 0, or 8 / 16 / 1008 with nx = 512 and ny a multiple of 8 / 16 / 8");
    const long long n = (long long)nimg * nx * ny;
    if (mode == 0 && (n % 4096 || (uintptr_t)src % 16 || (uintptr_t)dst % 16))
        return fail(MSX_EINVAL, "copy probe: mode 0 needs a multiple of 4096 values and 16-byte aligned buffers");
    HIPCHK(hipSetDevice(ctx->device));
    if (!ctx->ev0) {
        HIPCHK(hipEventCreate(&ctx->ev0));
        HIPCHK(hipEventCreate(&ctx->ev1));
    }
    const float2* s = static_cast<const float2*>(src);
    float2* d = static_cast<float2*>(dst);
    HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
    HIPCHK(mode == 0 ? launch_stream_copy(s, d, n, ctx->stream) : launch_tile_copy(s, d, nimg, ny, mode, ctx->stream));
    HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
    HIPCHK(hipEventSynchronize(ctx->ev1));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
    *ms = t;
    return MSX_OK;
}

MSX_API int msx_autotune(msx_ctx ctx, int nx, int ny, int prec, int max_batch, msx_tuning* best) {
    if (!ctx) return fail(MSX_ESTATE, "ctx is NULL");
    CHK(check_prec_where(prec, MSX_DEVICE));
    CHK(check_grid(nx, ny));
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (prec == MSX_F64) return autotune_impl<double2>(ctx, nx, ny, max_batch, best);
    return autotune_impl<float2>(ctx, nx, ny, max_batch, best);
}

MSX_API int msx_set_tuning(msx_ctx ctx, int nx, int ny, int prec, const msx_tuning* t) {
    if (!ctx) return fail(MSX_ESTATE, "ctx is NULL");
    CHK(check_prec_where(prec, MSX_DEVICE));
    CHK(check_grid(nx, ny));
    if (!t) {
        ctx->tune.erase(std::make_tuple(nx, ny, prec));
        return MSX_OK;
    }
    if (t->row_wpb < 0 || t->row_wpb > 64 || t->col_wpb < -8 || t->col_wpb > 64 || t->batch < 0)
        return fail(MSX_EINVAL, "bad tuning values");
    int tdef, talt;
    if (prec == MSX_F64) col_tiles<double2>(nx, &tdef, &talt);
    else col_tiles<float2>(nx, &tdef, &talt);
    if (t->col_tile != 0 && t->col_tile != tdef && t->col_tile != talt)
        return fail(MSX_EINVAL, "column tile %d: this grid takes 0, %d or %d", t->col_tile, tdef, talt);
    int ralt = 0, calt = 0, unused = 0;
    plan_alts(ny, &ralt, &unused);
    plan_alts(nx, &unused, &calt);
    if (t->row_plan < 0 || t->row_plan > ralt || t->col_plan < 0 || t->col_plan > calt)
        return fail(MSX_EINVAL, "FFT plans %d/%d: this grid takes row plan 0%s and column plan 0%s", t->row_plan,
                    t->col_plan, ralt ? " or 1" : "", calt ? " or 1" : "");
    ctx->tune[std::make_tuple(nx, ny, prec)] = *t;
    return MSX_OK;
}

MSX_API int msx_set_graphs(msx_ctx ctx, int enable) {
    if (!ctx) return fail(MSX_ESTATE, "ctx is NULL");
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (!enable) ctx->drop_graphs();
    ctx->graphs = enable != 0;
    return MSX_OK;
}

MSX_API int msx_graph_stats(msx_ctx ctx, int64_t* captures, int64_t* replays) {
    if (!ctx) return fail(MSX_ESTATE, "ctx is NULL");
    if (captures) *captures = ctx->graph_captures;
    if (replays) *replays = ctx->graph_hits;
    return MSX_OK;
}

MSX_API int msx_set_stream(msx_ctx ctx, void* s) {
    if (!ctx) return fail(MSX_ESTATE, "ctx is NULL");
    ctx->stream = s ? static_cast<hipStream_t>(s) : ctx->own;
    return MSX_OK;
}

MSX_API int msx_wait_stream(msx_ctx ctx, void* s) {
    if (!ctx) return fail(MSX_ESTATE, "ctx is NULL");
    HIPCHK(hipSetDevice(ctx->device));
    if (!ctx->ev_order) HIPCHK(hipEventCreateWithFlags(&ctx->ev_order, hipEventDisableTiming));
    HIPCHK(hipEventRecord(ctx->ev_order, static_cast<hipStream_t>(s)));
    HIPCHK(hipStreamWaitEvent(ctx->stream, ctx->ev_order, 0));
    return MSX_OK;
}

MSX_API double msx_wavelength(double kev) {
    const double emass = 510.99906, hc = 12.3984244;
    return hc / std::sqrt(kev * (2 * emass + kev));
}
MSX_API double msx_sigma(double kev) {
    const double emass = 510.99906;
    const double x = (emass + kev) / (2.0 * emass + kev);
    return 2.0 * (4.0 * std::atan(1.0)) * x / (msx_wavelength(kev) * kev);
}

static double h_bessi0(double x) {
    double ax = std::fabs(x);
    if (ax <= 3.75) {
        double t = x / 3.75;
        t = t * t;
        return 1.0 + t * (3.5156229 + t * (3.0899424 + t * (1.2067492 + t * (0.2659732 + t * (0.0360768 + t * 0.0045813)))));
    }
    double t = 3.75 / ax;
    double s = 0.39894228 + t * (0.01328592 + t * (0.00225319 + t * (-0.00157565 + t * (0.00916281 + t * (-0.02057706 +
               t * (0.02635537 + t * (-0.01647633 + t * 0.00392377)))))));
    return std::exp(ax) * s / std::sqrt(ax);
}
static double h_bessk0(double x) {
    double ax = std::fabs(x);
    if (ax > 0.0 && ax <= 2.0) {
        double x2 = ax / 2.0;
        x2 = x2 * x2;
        double s = -0.57721566 + x2 * (0.42278420 + x2 * (0.23069756 + x2 * (0.03488590 + x2 * (0.00262698 +
                   x2 * (0.00010750 + x2 * 0.00000740)))));
        return -std::log(ax / 2.0) * h_bessi0(x) + s;
    } else if (ax > 2.0) {
        double x2 = 2.0 / ax;
        double s = 1.25331414 + x2 * (-0.07832358 + x2 * (0.02189568 + x2 * (-0.01062446 + x2 * (0.00587872 +
                   x2 * (-0.00251540 + x2 * 0.00053208)))));
        return std::exp(-ax) * s / std::sqrt(ax);
    }
    return 1.0e20;
}

MSX_API double msx_vzatom(int Z, double r, int table, int formula) {
    if (Z < 1 || Z > 103 || table < 0 || table > 1) return 0.0;
    const double* p = MSX_FPARAMS[table][Z - 1];
    r = std::fabs(r);
    if (r < 1.0e-10) r = 1.0e-10;
    double suml = 0.0, sumg = 0.0;
    if (formula == MSX_FORMULA_KIRKLAND) {
        const double pi = 3.141592654;
        for (int i = 0; i < 6; i += 2) suml += p[i] * h_bessk0(2.0 * pi * r * std::sqrt(p[i + 1]));
        const double x2 = (pi * r) * (pi * r);
        for (int i = 6; i < 12; i += 2) sumg += p[i] * std::exp(-x2 / p[i + 1]) / p[i + 1];
        return 300.8242834 * suml + 150.4121417 * sumg;
    }
    for (int k = 0; k < 3; k++) suml += p[2 * k] * h_bessk0(6.2831853071796 * r * std::sqrt(p[2 * k + 1]));
    for (int k = 0; k < 3; k++) sumg += (p[6 + 2 * k] / p[7 + 2 * k]) * std::exp(-(6.2831853071796 * r * r) / p[7 + 2 * k]);
    return 300.73079394295 * suml + 150.36539697148 * sumg;
}
MSX_API double msx_featom(int Z, double k2, int table) {
    if (Z < 1 || Z > 103 || table < 0 || table > 1) return 0.0;
    const double* p = MSX_FPARAMS[table][Z - 1];
    double s = 0.0;
    for (int i = 0; i < 6; i += 2) s += p[i] / (k2 + p[i + 1]);
    for (int i = 6; i < 12; i += 2) s += p[i] * std::exp(-k2 * p[i + 1]);
    return s;
}

MSX_API int msx_displace(msx_ctx ctx, const msx_atom* in, int64_t natoms, const double* u_rms, uint64_t seed,
                         int32_t config, msx_atom* out, int where) {
    static_assert(sizeof(msx_atom) == sizeof(AtomIO), "atom layout");
    if (natoms < 0 || (natoms > 0 && (!in || !out)) || !u_rms) return fail(MSX_EINVAL, "bad atoms/u_rms");
    if (where == MSX_HOST) {
        const double twopi = 6.283185307179586476925287;
        const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
        for (int64_t i = 0; i < natoms; i++) {
            msx_atom a = in[i];
            const double u = (a.Z >= 0 && a.Z < 104) ? u_rms[a.Z] : 0.0;
            const uint32_t c0 = (uint32_t)i, c1 = (uint32_t)((uint64_t)i >> 32);
            const Philox4 r0 = philox4x32_10(c0, c1, (uint32_t)config, 0u, k0, k1);
            const Philox4 r1 = philox4x32_10(c0, c1, (uint32_t)config, 1u, k0, k1);
            const double u1 = uniform53(r0.v[0], r0.v[1]), u2 = uniform53(r0.v[2], r0.v[3]);
            const double u3 = uniform53(r1.v[0], r1.v[1]), u4 = uniform53(r1.v[2], r1.v[3]);
            const double ra = std::sqrt(-2.0 * std::log(u1)), rb = std::sqrt(-2.0 * std::log(u3));
            a.x += u * ra * std::cos(twopi * u2);
            a.y += u * ra * std::sin(twopi * u2);
            a.z += u * rb * std::cos(twopi * u4);
            out[i] = a;
        }
        return MSX_OK;
    }
    if (where != MSX_DEVICE) return fail(MSX_EINVAL, "where must be MSX_HOST or MSX_DEVICE");
    if (!ctx) return fail(MSX_ESTATE, "ctx is NULL");
    HIPCHK(hipSetDevice(ctx->device));
    CHK(ctx->aux.ensure(104 * sizeof(double)));
    HIPCHK(hipMemcpyAsync(ctx->aux.p, u_rms, 104 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(launch_displace(reinterpret_cast<const AtomIO*>(in), natoms, static_cast<const double*>(ctx->aux.p), seed,
                           config, reinterpret_cast<AtomIO*>(out), ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return MSX_OK;
}

// Philox-4x32-10 block for the known-answer tests
MSX_API void msx_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    const Philox4 r = philox4x32_10(ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1]);
    for (int i = 0; i < 4; i++) out[i] = r.v[i];
}

MSX_API int msx_fft2d(msx_ctx ctx, void* data, int nbatch, int nx, int ny, int sign, int prec, int where) {
    if (!ctx) return fail(MSX_ESTATE, "ctx is NULL");
    CHK(check_prec_where(prec, where));
    CHK(check_grid(nx, ny));
    if (!data || nbatch <= 0 || (sign != 1 && sign != -1)) return fail(MSX_EINVAL, "bad data/nbatch/sign");
    HIPCHK(hipSetDevice(ctx->device));
    const size_t es = prec == MSX_F64 ? 16 : 8, bytes = es * (size_t)nx * ny * nbatch;
    void* d = data;
    if (where == MSX_HOST) {
        CHK(ctx->wave.ensure(bytes));
        d = ctx->wave.p;
        HIPCHK(hipMemcpyAsync(d, data, bytes, hipMemcpyHostToDevice, ctx->stream));
    }
    if (prec == MSX_F64) CHK(fft2d_dev<double2>(ctx, static_cast<double2*>(d), nbatch, nx, ny, sign));
    else CHK(fft2d_dev<float2>(ctx, static_cast<float2*>(d), nbatch, nx, ny, sign));
    if (where == MSX_HOST) HIPCHK(hipMemcpyAsync(data, d, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return MSX_OK;
}

MSX_API int msx_potential(msx_ctx ctx, const msx_atom* atoms, int64_t natoms, double a, double b, double c, int nx,
                          int ny, int nslices, double radius, int table, int formula, void* V, int prec, int where) {
    if (!ctx) return fail(MSX_ESTATE, "ctx is NULL");
    CHK(check_prec_where(prec, where));
    if (nx <= 0 || ny <= 0 || nx > 16384 || ny > 16384 || nslices <= 0) return fail(MSX_EINVAL, "bad grid %d x %d x %d", nx, ny, nslices);
    if (!(a > 0) || !(b > 0) || !(c > 0) || !(radius > 0)) return fail(MSX_EINVAL, "a, b, c, radius must be > 0");
    if (table < 0 || table > 1 || formula < 0 || formula > 1) return fail(MSX_EINVAL, "bad table/formula");
    if (natoms < 0 || (natoms > 0 && !atoms) || !V) return fail(MSX_EINVAL, "bad atoms/V");
    HIPCHK(hipSetDevice(ctx->device));
    return potential_impl(ctx, atoms, natoms, a, b, c, nx, ny, nslices, radius, table, formula, V, prec, where);
}

MSX_API int msx_potential_device_atoms(msx_ctx ctx, const msx_atom* atoms, int64_t natoms, double a, double b,
                                       double c, int nx, int ny, int nslices, double radius, int table, int formula,
                                       void* V, int prec, int where) {
    if (!ctx) return fail(MSX_ESTATE, "ctx is NULL");
    CHK(check_prec_where(prec, where));
    if (nx <= 0 || ny <= 0 || nx > 16384 || ny > 16384 || nslices <= 0) return fail(MSX_EINVAL, "bad grid %d x %d x %d", nx, ny, nslices);
    if (!(a > 0) || !(b > 0) || !(c > 0) || !(radius > 0)) return fail(MSX_EINVAL, "a, b, c, radius must be > 0");
    if (table < 0 || table > 1 || formula < 0 || formula > 1) return fail(MSX_EINVAL, "bad table/formula");
    if (natoms < 0 || (natoms > 0 && !atoms) || !V) return fail(MSX_EINVAL, "bad atoms/V");
    HIPCHK(hipSetDevice(ctx->device));
    return potential_dev_impl(ctx, atoms, natoms, a, b, c, nx, ny, nslices, radius, table, formula, V, prec, where);
}

MSX_API int msx_potential_rec(msx_ctx ctx, const msx_atom* atoms, int64_t natoms, const double* dw,
                              const double* ads, double a, double b, double c, double gam, int nmult, int mmult,
                              int nx, int ny, int nslices, void* V, void* Vabs, int prec, int where) {
    if (!ctx) return fail(MSX_ESTATE, "ctx is NULL");
    CHK(check_prec_where(prec, where));
    CHK(check_grid(nx, ny));
    if (nslices <= 0) return fail(MSX_EINVAL, "nslices must be > 0");
    if (!(a > 0) || !(b > 0) || !(c > 0)) return fail(MSX_EINVAL, "a, b, c must be > 0");
    if (!(gam > 0.0 && gam < 3.14159265358979323846)) return fail(MSX_EINVAL, "gam must be in (0, pi)");
    if (nmult < 1 || mmult < 1) return fail(MSX_EINVAL, "nmult and mmult must be >= 1");
    if (natoms < 0 || (natoms > 0 && !atoms) || !V) return fail(MSX_EINVAL, "bad atoms/V");
    if (natoms > (1LL << 30)) return fail(MSX_EINVAL, "too many atoms");
    HIPCHK(hipSetDevice(ctx->device));
    return potential_rec_impl(ctx, atoms, natoms, dw, ads, a, b, c, gam, nmult, mmult, nx, ny, nslices, V, Vabs, prec,
                              where);
}

MSX_API int msx_potential_rs(msx_ctx ctx, const msx_atom* atoms, int64_t natoms, const double* dw,
                             const double* ads, double a, double b, double c, double gam, double atrad, int periodic,
                             int nx, int ny, int nslices, void* V, void* Vabs, int prec, int where) {
    if (!ctx) return fail(MSX_ESTATE, "ctx is NULL");
    CHK(check_prec_where(prec, where));
    if (nx < 1 || ny < 1 || nx > 16384 || ny > 16384 || nslices <= 0)
        return fail(MSX_EINVAL, "bad grid %d x %d x %d", nx, ny, nslices);
    if (!(a > 0) || !(b > 0) || !(c > 0) || !(atrad > 0)) return fail(MSX_EINVAL, "a, b, c, atrad must be > 0");
    if (!(gam > 0.0 && gam < 3.14159265358979323846)) return fail(MSX_EINVAL, "gam must be in (0, pi)");
    if (natoms < 0 || (natoms > 0 && !atoms) || !V) return fail(MSX_EINVAL, "bad atoms/V");
    if (natoms > (1LL << 30)) return fail(MSX_EINVAL, "too many atoms");
    HIPCHK(hipSetDevice(ctx->device));
    return potential_rs_impl(ctx, atoms, natoms, dw, ads, a, b, c, gam, atrad, periodic, nx, ny, nslices, V, Vabs,
                             prec, where);
}

MSX_API int msx_create_pots_rec_dt(int ntypes, int npoints, double dds, const int32_t* types, const double* dw,
                                   double vol, double* sc) {
    if (ntypes < 1 || npoints < 1 || !types || !dw || !sc || !(dds > 0))
        return fail(MSX_EINVAL, "curves: ntypes, npoints >= 1, dds > 0, types, dw and sc required");
    for (int nt = 0; nt < ntypes; nt++)
        if (types[nt] < 0 || types[nt] > 92) return fail(MSX_EINVAL, "type %d: Z = %d outside 0..92", nt, types[nt]);
    for (int nt = 0; nt < ntypes; nt++) {
        const double* ab = MSX_DOYLE_TURNER[types[nt]];
        double s = 0.0;
        for (int i = 0; i < npoints; i++, s += dds) {
            const double s2 = s * s;
            double f = 0.0;
            for (int j = 0; j < 4; j++) f += ab[2 * j] * std::exp(-ab[2 * j + 1] * s2);
            sc[(size_t)nt * npoints + i] = f * std::exp(-dw[nt] * s * s) * vol;
        }
    }
    return MSX_OK;
}

MSX_API int msx_create_pots_rs_dt(int ntypes, int npoints, double dds, const int32_t* types, const double* dw,
                                  double* sc) {
    if (ntypes < 1 || npoints < 1 || !types || !dw || !sc || !(dds > 0))
        return fail(MSX_EINVAL, "curves: ntypes, npoints >= 1, dds > 0, types, dw and sc required");
    for (int nt = 0; nt < ntypes; nt++)
        if (types[nt] < 0 || types[nt] > 92) return fail(MSX_EINVAL, "type %d: Z = %d outside 0..92", nt, types[nt]);
    for (int nt = 0; nt < ntypes; nt++) {
        const double* ab = MSX_DOYLE_TURNER[types[nt]];
        double s = 0.0;
        for (int i = 0; i < npoints; i++, s += dds) {
            const double e = -4.0 * M_PI * M_PI * s * s;
            double f = 0.0;
            for (int j = 0; j < 4; j++)
                if (ab[2 * j + 1] > 0) f += ab[2 * j] / (ab[2 * j + 1] + dw[nt]) * std::exp(e / (ab[2 * j + 1] + dw[nt]));
            sc[(size_t)nt * npoints + i] = f * 4 * M_PI;
        }
    }
    return MSX_OK;
}

MSX_API int msx_potential_rec_curve(msx_ctx ctx, const msx_atom* atoms, int64_t natoms, const int32_t* type,
                                    const double* curves, int ntypes, int npoints, double dds, const double* ads,
                                    double a, double b, double c, double gam, int nmult, int mmult, int nx, int ny,
                                    int nslices, void* V, void* Vabs, int prec, int where) {
    if (!ctx) return fail(MSX_ESTATE, "ctx is NULL");
    CHK(check_prec_where(prec, where));
    CHK(check_grid(nx, ny));
    if (nslices <= 0 || !(a > 0) || !(b > 0) || !(c > 0) || !(gam > 0.0 && gam < 3.14159265358979323846) ||
        nmult < 1 || mmult < 1)
        return fail(MSX_EINVAL, "bad cell, slices or multipliers");
    if (natoms < 0 || (natoms > 0 && !atoms) || !V || natoms > (1LL << 30)) return fail(MSX_EINVAL, "bad atoms/V");
    const CurveSpec cs{type, curves, ntypes, npoints, dds};
    CHK(check_curves(&cs, natoms));
    HIPCHK(hipSetDevice(ctx->device));
    return potential_rec_impl(ctx, atoms, natoms, nullptr, ads, a, b, c, gam, nmult, mmult, nx, ny, nslices, V, Vabs,
                              prec, where, &cs);
}

MSX_API int msx_potential_rs_curve(msx_ctx ctx, const msx_atom* atoms, int64_t natoms, const int32_t* type,
                                   const double* curves, int ntypes, int npoints, double dds, const double* ads,
                                   double a, double b, double c, double gam, double atrad, int periodic, int nx,
                                   int ny, int nslices, void* V, void* Vabs, int prec, int where) {
    if (!ctx) return fail(MSX_ESTATE, "ctx is NULL");
    CHK(check_prec_where(prec, where));
    if (nx < 1 || ny < 1 || nx > 16384 || ny > 16384 || nslices <= 0)
        return fail(MSX_EINVAL, "bad grid %d x %d x %d", nx, ny, nslices);
    if (!(a > 0) || !(b > 0) || !(c > 0) || !(atrad > 0) || !(gam > 0.0 && gam < 3.14159265358979323846))
        return fail(MSX_EINVAL, "bad cell or atrad");
    if (natoms < 0 || (natoms > 0 && !atoms) || !V || natoms > (1LL << 30)) return fail(MSX_EINVAL, "bad atoms/V");
    const CurveSpec cs{type, curves, ntypes, npoints, dds};
    CHK(check_curves(&cs, natoms));
    HIPCHK(hipSetDevice(ctx->device));
    return potential_rs_impl(ctx, atoms, natoms, nullptr, ads, a, b, c, gam, atrad, periodic, nx, ny, nslices, V, Vabs,
                             prec, where, &cs);
}

MSX_API int msx_transmission(msx_ctx ctx, const void* V, int nlayer, int nx, int ny, double ax, double by, double kev,
                             double bw, void* t, int prec, int where) {
    return msx_transmission_abs(ctx, V, nullptr, nlayer, nx, ny, ax, by, kev, bw, t, prec, where);
}

MSX_API int msx_transmission_abs(msx_ctx ctx, const void* V, const void* Vabs, int nlayer, int nx, int ny, double ax,
                                 double by, double kev, double bw, void* t, int prec, int where) {
    if (!ctx) return fail(MSX_ESTATE, "ctx is NULL");
    CHK(check_prec_where(prec, where));
    CHK(check_grid(nx, ny));
    if (!V || !t || nlayer <= 0 || !(ax > 0) || !(by > 0) || !(kev > 0)) return fail(MSX_EINVAL, "bad arguments");
    HIPCHK(hipSetDevice(ctx->device));
    const size_t n = (size_t)nx * ny * nlayer;
    const size_t vb = (prec == MSX_F64 ? 8 : 4) * n, tb = (prec == MSX_F64 ? 16 : 8) * n;
    const void* v_dev = V;
    const void* a_dev = Vabs;
    void* t_dev = t;
    if (where == MSX_HOST) {
        CHK(ctx->vtmp.ensure(vb * (Vabs ? 2 : 1)));
        CHK(ctx->trans.ensure(tb));
        HIPCHK(hipMemcpyAsync(ctx->vtmp.p, V, vb, hipMemcpyHostToDevice, ctx->stream));
        v_dev = ctx->vtmp.p;
        if (Vabs) {
            a_dev = static_cast<char*>(ctx->vtmp.p) + vb;
            HIPCHK(hipMemcpyAsync(const_cast<void*>(a_dev), Vabs, vb, hipMemcpyHostToDevice, ctx->stream));
        }
        t_dev = ctx->trans.p;
    }
    if (prec == MSX_F64)
        CHK(transmission_dev<double2>(ctx, static_cast<const double*>(v_dev), nlayer, nx, ny, ax, by, kev, bw,
                                      static_cast<double2*>(t_dev), static_cast<const double*>(a_dev)));
    else
        CHK(transmission_dev<float2>(ctx, static_cast<const float*>(v_dev), nlayer, nx, ny, ax, by, kev, bw,
                                     static_cast<float2*>(t_dev), static_cast<const float*>(a_dev)));
    if (where == MSX_HOST) HIPCHK(hipMemcpyAsync(t, t_dev, tb, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return MSX_OK;
}

MSX_API int msx_multislice_batch(msx_ctx ctx, const msx_job* jobs, int njobs, int prec, int where) {
    if (!ctx) return fail(MSX_ESTATE, "ctx is NULL");
    CHK(check_prec_where(prec, where));
    if (where != MSX_DEVICE) return fail(MSX_EINVAL, "msx_multislice_batch takes device buffers (MSX_DEVICE)");
    if (njobs < 0 || (njobs > 0 && !jobs)) return fail(MSX_EINVAL, "bad jobs");
    for (int i = 0; i < njobs; i++) {
        if (!jobs[i].stack || !jobs[i].trans || !jobs[i].psi) return fail(MSX_EINVAL, "job %d: stack, trans, psi required", i);
        if (is_oblique(jobs[i].stack))
            return fail(MSX_EINVAL, "job %d: oblique cell (gam != pi/2): run it through msx_multislice", i);
    }
    HIPCHK(hipSetDevice(ctx->device));
    if (prec == MSX_F64) return multislice_batch_impl<double2>(ctx, jobs, njobs);
    return multislice_batch_impl<float2>(ctx, jobs, njobs);
}

MSX_API int msx_multislice(msx_ctx ctx, const msx_stack* stack, const void* trans, const void* psi0, void* psi,
                           int nwave, int prec, int where) {
    if (!ctx) return fail(MSX_ESTATE, "ctx is NULL");
    CHK(check_prec_where(prec, where));
    CHK(check_stack(stack));
    if (!trans || !psi || nwave <= 0) return fail(MSX_EINVAL, "bad trans/psi/nwave");
    HIPCHK(hipSetDevice(ctx->device));
    if (prec == MSX_F64) return multislice_impl<double2>(ctx, stack, trans, psi0, psi, nwave, where);
    return multislice_impl<float2>(ctx, stack, trans, psi0, psi, nwave, where);
}

MSX_API int msx_multislice_series(msx_ctx ctx, const msx_stack* stack, const void* trans, const void* psi0,
                                  void* psi, int nwave, int prec, int where, const msx_series* series) {
    if (!ctx) return fail(MSX_ESTATE, "ctx is NULL");
    CHK(check_prec_where(prec, where));
    CHK(check_stack(stack));
    if (!trans || !psi || nwave <= 0) return fail(MSX_EINVAL, "bad trans/psi/nwave");
    if (nwave > 65535) return fail(MSX_EINVAL, "msx_multislice_series: at most 65535 waves (%d)", nwave);
    if (!series) return fail(MSX_EINVAL, "msx_multislice_series: series is NULL");
    const msx_series* se = series;
    if (se->nat < 0 || se->nbeams < 0 || se->nbeams > 4096)
        return fail(MSX_EINVAL, "msx_multislice_series: nat %d, nbeams %d (0..4096)", se->nat, se->nbeams);
    if (se->nat > 0) {
        if (!se->at || !se->waves) return fail(MSX_EINVAL, "msx_multislice_series: nat > 0 needs at and waves");
        for (int i = 0; i < se->nat; i++) {
            const int a = se->at[i];
            if (a < 1 || a > stack->nslice || (i > 0 && a <= se->at[i - 1]))
                return fail(MSX_EINVAL,
                            "msx_multislice_series: at[%d] = %d: slice counts must ascend strictly within 1..%d", i,
                            a, stack->nslice);
        }
    }
    if (se->nbeams > 0 && (!se->hk || !se->beams))
        return fail(MSX_EINVAL, "msx_multislice_series: nbeams > 0 needs hk and beams");
    HIPCHK(hipSetDevice(ctx->device));
    if (prec == MSX_F64) return multislice_series_impl<double2>(ctx, stack, trans, psi0, psi, nwave, where, se);
    return multislice_series_impl<float2>(ctx, stack, trans, psi0, psi, nwave, where, se);
}

MSX_API int msx_multislice_noalloc(msx_ctx ctx, const msx_stack* stack, const void* trans, const void* psi0,
                                   void* psi, int nwave, int prec) {
    if (!ctx) return fail(MSX_ESTATE, "ctx is NULL");
    CHK(check_prec_where(prec, MSX_DEVICE));
    CHK(check_stack(stack));
    if (!trans || !psi || nwave <= 0) return fail(MSX_EINVAL, "bad trans/psi/nwave");
    HIPCHK(hipSetDevice(ctx->device));
    if (prec == MSX_F64) return multislice_noalloc_impl<double2>(ctx, stack, trans, psi0, psi, nwave);
    return multislice_noalloc_impl<float2>(ctx, stack, trans, psi0, psi, nwave);
}

MSX_API int msx_stem(msx_ctx ctx, const msx_stack* stack, const void* trans, const msx_probe* probe, const double* pos,
                     int nprobe, const msx_detector* det, int ndet, const int32_t* ck, int nck, double weight,
                     int accumulate, double* out, int batch, int prec, int where) {
    if (!ctx) return fail(MSX_ESTATE, "ctx is NULL");
    CHK(check_prec_where(prec, where));
    CHK(check_stack(stack));
    if (!trans || !probe || !pos || nprobe <= 0 || !out) return fail(MSX_EINVAL, "bad trans/probe/pos/out");
    if (ndet <= 0 || ndet > kMaxDet || !det) return fail(MSX_EINVAL, "ndet must be 1..%d", kMaxDet);
    if (nck <= 0 || !ck) return fail(MSX_EINVAL, "need at least one checkpoint");
    for (int i = 0; i < nck; i++)
        if (ck[i] < 1 || ck[i] > stack->nslice || (i && ck[i] <= ck[i - 1]))
            return fail(MSX_EINVAL, "checkpoints must increase within 1..nslice");
    if (!(probe->alpha_max > 0)) return fail(MSX_EINVAL, "alpha_max must be > 0");
    if (is_oblique(stack))
        return fail(MSX_EINVAL, "msx_stem: oblique cells (gam != pi/2) are not supported (the probe and the detectors "
                                "are built on an orthogonal cell)");
    HIPCHK(hipSetDevice(ctx->device));
    if (prec == MSX_F64)
        return stem_impl<double2>(ctx, stack, trans, probe, pos, nprobe, det, ndet, ck, nck, weight, accumulate, out,
                                  batch, where);
    return stem_impl<float2>(ctx, stack, trans, probe, pos, nprobe, det, ndet, ck, nck, weight, accumulate, out, batch,
                             where);
}

MSX_API int msx_partial_coherence(msx_ctx ctx, const msx_stack* stack, const void* trans, const msx_lens* lens,
                                  double acmin, double acmax, double sigmaf, double dfdelt, int batch, double* image,
                                  int32_t* nangles, int32_t* nsteps, int prec, int where) {
    if (!ctx) return fail(MSX_ESTATE, "ctx is NULL");
    CHK(check_prec_where(prec, where));
    CHK(check_stack(stack));
    if (!trans || !lens || !image) return fail(MSX_EINVAL, "trans, lens and image must be set");
    if (stack->nx < 2 || stack->ny < 2) return fail(MSX_ESIZE, "partial coherence needs nx, ny >= 2");
    if (!(acmin >= 0) || !(acmax >= acmin)) return fail(MSX_EINVAL, "need 0 <= acmin <= acmax");
    if (!(lens->aobj >= 0)) return fail(MSX_EINVAL, "aobj must be >= 0");
    if (!(sigmaf >= 0)) return fail(MSX_EINVAL, "sigmaf must be >= 0");
    if (sigmaf > 0 && !(dfdelt > 0)) return fail(MSX_EINVAL, "dfdelt must be > 0 when sigmaf > 0");
    if (sigmaf > 0 && (2.5 * sigmaf / dfdelt) > 4096) return fail(MSX_EINVAL, "more than 8193 defocus steps");
    if (batch < 0) return fail(MSX_EINVAL, "batch must be >= 0");
    if (is_oblique(stack)) return fail(MSX_EINVAL, "msx_partial_coherence: oblique cells are not supported");
    HIPCHK(hipSetDevice(ctx->device));
    if (prec == MSX_F64)
        return partial_impl<double2>(ctx, stack, trans, lens, acmin, acmax, sigmaf, dfdelt, batch, image, nangles,
                                     nsteps, where);
    return partial_impl<float2>(ctx, stack, trans, lens, acmin, acmax, sigmaf, dfdelt, batch, image, nangles, nsteps,
                                where);
}

MSX_API int msx_ctf_image(msx_ctx ctx, const void* psi, int nwave, int nx, int ny, double ax, double by, double kev,
                          double cs, double df, double aobj, void* psi_img, void* intensity, int prec, int where) {
    if (!ctx) return fail(MSX_ESTATE, "ctx is NULL");
    CHK(check_prec_where(prec, where));
    CHK(check_grid(nx, ny));
    if (!psi || nwave <= 0 || !(kev > 0)) return fail(MSX_EINVAL, "bad arguments");
    HIPCHK(hipSetDevice(ctx->device));
    if (prec == MSX_F64)
        return ctf_impl<double2>(ctx, psi, nwave, nx, ny, ax, by, kev, cs, df, aobj, psi_img, intensity, where);
    return ctf_impl<float2>(ctx, psi, nwave, nx, ny, ax, by, kev, cs, df, aobj, psi_img, intensity, where);
}

}  // extern "C"
