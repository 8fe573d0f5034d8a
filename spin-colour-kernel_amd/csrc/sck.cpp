// libsck: the C-ABI of include/sck.h.
//
// Compiles the HIP source generated per system with hipRTC for gfx950 (cached as code objects on disk, keyed
// by a hash of source + options and verified against the stored source), loads the module, resolves its
// kernels, and executes the launch plans the host generator computed: coloured sweeps, the sequential chain,
// the synchronous control, the float64 observables, and HIP graphs of k sweeps.
#include "sck.h"

#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <cstdio>
#include <cstring>
#include <mutex>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>
#include <sys/stat.h>
#include <unistd.h>

// static kernels of the learned-proposal training (mlp.hip)
extern "C" __global__ void sck_mlp_fwdbwd_k(const float *, const sck_mlp_sample *, int32_t, sck_mlp_config, float *);
extern "C" __global__ void sck_mlp_reduce_k(const float *, int32_t, int32_t, float *);
extern "C" __global__ void sck_mlp_adam_k(float *, float *, float *, float *, const float *, int32_t,
                                        sck_mlp_adam_config);
extern "C" __global__ void sck_mlp_table_k(const float *, sck_mlp_config, sck_width_grid, float *);

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                                   \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            return fail(SCK_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));                   \
    } while (0)

const char *kArch = "gfx950";

std::vector<std::string> compile_options() {
    return {"--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
            // packed fp32 instructions are left to the compiler's scalar path (no v_pk_*_f32)
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
}

uint64_t fnv1a(const std::string &s, uint64_t h = 1469598103934665603ull) {
    for (unsigned char c : s) {
        h ^= c;
        h *= 1099511628211ull;
    }
    return h;
}

std::string cache_key(const std::string &src) {
    std::string all = src;
    for (auto &o : compile_options()) all += "\n//" + o;
    all += "\n//arch=";
    all += kArch;
    char buf[64];
    std::snprintf(buf, sizeof buf, "%016llx%08zx", (unsigned long long)fnv1a(all), all.size());
    return buf;
}

bool read_file(const std::string &path, std::string &out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::stringstream ss;
    ss << f.rdbuf();
    out = ss.str();
    return true;
}

bool write_file_atomic(const std::string &path, const std::string &data) {
    // a temporary name of this process's own, so that processes filling one cache never share a file
    std::string tmp = path + ".tmp" + std::to_string((long long)::getpid());
    {
        std::ofstream f(tmp, std::ios::binary);
        if (!f) return false;
        f.write(data.data(), (std::streamsize)data.size());
        if (!f) return false;
    }
    return std::rename(tmp.c_str(), path.c_str()) == 0;
}

void mkdirs(const std::string &dir) {
    std::string cur;
    for (size_t i = 0; i < dir.size(); ++i) {
        cur += dir[i];
        if (dir[i] == '/' || i + 1 == dir.size()) ::mkdir(cur.c_str(), 0755);
    }
}

int rtc_compile(const std::string &src, std::string &code) {
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "sck_system.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS)
        return fail(SCK_E_COMPILE, "hiprtcCreateProgram failed");
    auto opts = compile_options();
    std::vector<const char *> argv;
    for (auto &o : opts) argv.push_back(o.c_str());
    hiprtcResult r = hiprtcCompileProgram(prog, (int)argv.size(), argv.data());
    size_t logsz = 0;
    hiprtcGetProgramLogSize(prog, &logsz);
    std::string log(logsz, '\0');
    if (logsz) hiprtcGetProgramLog(prog, &log[0]);
    if (r != HIPRTC_SUCCESS) {
        hiprtcDestroyProgram(&prog);
        return fail(SCK_E_COMPILE, std::string("hipRTC compilation failed: ") + hiprtcGetErrorString(r) + "\n" + log);
    }
    size_t sz = 0;
    hiprtcGetCodeSize(prog, &sz);
    code.assign(sz, '\0');
    hiprtcGetCode(prog, &code[0]);
    hiprtcDestroyProgram(&prog);
    return SCK_OK;
}

// code object for `src`: from the cache directory when present (source verified), else compiled and stored
int code_object(const std::string &src, const char *cache_dir, std::string &code) {
    std::string key = cache_key(src);
    std::string dir = cache_dir ? cache_dir : "";
    if (!dir.empty()) {
        std::string cached_src;
        if (read_file(dir + "/" + key + ".hip", cached_src) && cached_src == src &&
            read_file(dir + "/" + key + ".co", code) && !code.empty())
            return SCK_OK;
    }
    int rc = rtc_compile(src, code);
    if (rc) return rc;
    if (!dir.empty()) {
        mkdirs(dir);
        write_file_atomic(dir + "/" + key + ".co", code);
        write_file_atomic(dir + "/" + key + ".hip", src);
    }
    return SCK_OK;
}

}  // namespace

struct sck_handle {
    hipModule_t module = nullptr;
    std::vector<hipFunction_t> fn;
    std::vector<std::string> names;
    std::vector<sck_launch> sweep, ref, sync, obs;
    sck_launch bump{};
    bool has_resident = false;
    sck_launch resident{};
    bool has_tiled = false;
    sck_launch tiled{};
    bool has_fold = false;
    sck_launch fold{};        // running totals of E and M: adds one tiled sweep's partial sums
    size_t spin_bytes = 0;    // float32 spins of the whole batch (the tiled sweep's buffers)
    int sample_kernel = -1;
    int mode = 0, n = 0, b = 0, batch = 0, words = 0, obs_chunks = 1;
    int rep_blocks = 1;         // the sweep list holds rep_blocks groups of class launches, one per replica block
    int obs_blocks = 1;         // the observation list holds obs_blocks groups (1, or rep_blocks: one per block)
};

struct sck_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t stream = nullptr;
};

namespace {

int launch(sck_handle *h, const sck_launch &l, const sck_args &a, hipStream_t s) {
    if (l.kernel < 0 || l.kernel >= (int)h->fn.size()) return fail(SCK_E_INVALID, "launch plan names an unknown kernel");
    sck_args copy = a;
    copy.rep_base = l.rep_base;
    if (l.tiles) {
        copy.tiles = l.tiles;
        copy.ntile = l.ntile;
    }
    void *params[] = {&copy};
    HIP_TRY(hipModuleLaunchKernel(h->fn[l.kernel], l.grid[0], l.grid[1], l.grid[2], l.block[0], l.block[1],
                                  l.block[2], l.shmem, s, params, nullptr));
    return SCK_OK;
}

int bump(sck_handle *h, const sck_args &a, int64_t by, hipStream_t s) {
    sck_args b = a;
    b.sweep_offset = by;
    return launch(h, h->bump, b, s);
}

// Event pairs (start, stop, start, stop, ...) recorded on the launch stream around sweep-kernel launches; the
// counter bump stays outside them. Destroyed with the timer (a pending event may be destroyed).
struct launch_timer {
    std::vector<hipEvent_t> ev;
    ~launch_timer() {
        for (hipEvent_t e : ev)
            if (e) hipEventDestroy(e);
    }
};

int timed_launch(sck_handle *h, const sck_launch &l, const sck_args &a, hipStream_t s, launch_timer *t) {
    if (!t) return launch(h, l, a, s);
    hipEvent_t e[2] = {nullptr, nullptr};
    for (hipEvent_t &x : e) {
        HIP_TRY(hipEventCreate(&x));
        t->ev.push_back(x);
    }
    HIP_TRY(hipEventRecord(e[0], s));
    int rc = launch(h, l, a, s);
    if (rc) return rc;
    HIP_TRY(hipEventRecord(e[1], s));
    return SCK_OK;
}

// Whether n sweeps over classes [lo, hi) run through the tiled kernel (whole sweeps with a workspace).
bool tiled_sweeps(const sck_handle *h, const sck_args *args, int32_t n, int32_t lo, int32_t hi) {
    const int nb = h->rep_blocks, ns = (int)h->sweep.size() / nb;
    if (hi < 0 || hi > ns) hi = ns;
    return h->has_tiled && !h->has_resident && lo == 0 && hi == ns && nb == 1 && args->dst && n > 0;
}

int do_sweeps(sck_handle *h, const sck_args *args, int32_t n, int32_t lo, int32_t hi, hipStream_t s,
              launch_timer *t = nullptr, bool copy_back = true) {
    // lo, hi: class range; the sweep list is rep_blocks groups of ns class launches
    const int nb = h->rep_blocks, ns = (int)h->sweep.size() / nb;
    if (hi < 0 || hi > ns) hi = ns;
    if (lo < 0 || lo > hi) return fail(SCK_E_INVALID, "class range out of bounds");
    sck_args a = *args;
    const bool tiled_path = tiled_sweeps(h, args, n, lo, hi);
    if (args->track && n > 0 && !(tiled_path && h->has_fold && args->track_parts))
        return fail(SCK_E_INVALID, "running totals of E and M need whole tiled sweeps (a module with a tiled "
                                   "kernel, args.dst and args.track_parts)");
    if (h->has_resident && lo == 0 && hi == ns && n > 0) {
        // whole sweeps: one launch, each workgroup runs its replica's n sweeps with the lattice in LDS
        a.sweep_offset = 0;
        a.nsweep = n;
        int rc = timed_launch(h, h->resident, a, s, t);
        if (rc) return rc;
        return bump(h, *args, n, s);
    }
    if (tiled_path) {
        // whole sweeps: one launch per sweep, every class of a (replica, tile) in LDS. A tile reads its halo
        // from other tiles' cells, so a sweep reads one buffer and writes the other (args->dst, a workspace of
        // the spins' size); after an odd number of sweeps the result is copied back into args->spins
        for (int32_t i = 0; i < n; ++i) {
            a.sweep_offset = i;
            a.spins = (i & 1) ? args->dst : args->spins;
            a.dst = (i & 1) ? args->spins : args->dst;
            int rc = timed_launch(h, h->tiled, a, s, t);
            if (rc) return rc;
            if (args->track) {
                rc = launch(h, h->fold, a, s);
                if (rc) return rc;
            }
        }
        // an odd count leaves the result in args->dst: copied back, unless the caller keeps it there
        if ((n & 1) && copy_back)
            HIP_TRY(hipMemcpyAsync(args->spins, args->dst, h->spin_bytes, hipMemcpyDeviceToDevice, s));
        return bump(h, *args, n, s);
    }
    // each replica block runs all n sweeps before the next one starts
    for (int rb = 0; rb < nb; ++rb) {
        for (int32_t i = 0; i < n; ++i) {
            a.sweep_offset = i;
            for (int k = lo; k < hi; ++k) {
                int rc = timed_launch(h, h->sweep[(size_t)rb * ns + k], a, s, t);
                if (rc) return rc;
            }
        }
    }
    return n > 0 ? bump(h, *args, n, s) : SCK_OK;
}

int do_observe(sck_handle *h, const sck_args *args, hipStream_t s) {
    for (auto &l : h->obs) {
        int rc = launch(h, l, *args, s);
        if (rc) return rc;
    }
    return SCK_OK;
}

// k whole sweeps, then an observation. With one observation group per replica block, each block's group runs
// right after that block's k sweeps, while the block's lattice is still in the cache; the blocks are
// independent chains, so the values equal those of observing after every block has swept.
int do_sweeps_observed(sck_handle *h, const sck_args *args, int32_t k, hipStream_t s) {
    const int nb = h->rep_blocks;
    if (h->obs_blocks <= 1 || h->obs_blocks != nb || h->has_resident || h->has_tiled || k <= 0) {
        int rc = do_sweeps(h, args, k, 0, -1, s);
        return rc ? rc : do_observe(h, args, s);
    }
    if (args->track) return fail(SCK_E_INVALID, "running totals of E and M need whole tiled sweeps");
    const int ns = (int)h->sweep.size() / nb, no = (int)h->obs.size() / nb;
    sck_args a = *args;
    for (int rb = 0; rb < nb; ++rb) {
        for (int32_t i = 0; i < k; ++i) {
            a.sweep_offset = i;
            for (int c = 0; c < ns; ++c) {
                int rc = launch(h, h->sweep[(size_t)rb * ns + c], a, s);
                if (rc) return rc;
            }
        }
        for (int j = 0; j < no; ++j) {
            int rc = launch(h, h->obs[(size_t)rb * no + j], *args, s);
            if (rc) return rc;
        }
    }
    return bump(h, *args, k, s);
}

// The partial sums of sck_observe_noalloc: a file-scope device array, one copy per GPU, which holds its
// contents between calls (every call on a GPU shares it; calls that overlap on one GPU race). The host-side name is
// only a handle: its device address is read once per device with hipGetSymbolAddress and used from then on.
static __device__ __align__(256) double scratch[SCK_OBS_SCRATCH_BYTES / sizeof(double)];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

int obs_scratch_ptr(double **out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(SCK_E_INVALID, "device index beyond the library's table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = (double *)p;
    }
    *out = g_scratch_ptr[dev];
    return SCK_OK;
}

}  // namespace

extern "C" {

int sck_version(void) { return SCK_ABI_VERSION; }

const char *sck_last_error(void) { return g_err.c_str(); }

int sck_precompile(const char *hip_src, const char *cache_dir) {
    if (!hip_src) return fail(SCK_E_INVALID, "null source");
    std::string code;
    return code_object(hip_src, cache_dir, code);
}

int sck_create(const sck_desc *d, const char *hip_src, const char *cache_dir, sck_handle **out) {
    if (!d || !hip_src || !out) return fail(SCK_E_INVALID, "null argument");
    if (d->abi_version != SCK_ABI_VERSION) return fail(SCK_E_INVALID, "descriptor ABI version mismatch");
    if (d->n_kernels <= 0 || !d->kernel_names) return fail(SCK_E_INVALID, "no kernels");
    if (!d->bump || d->n_sweep < 0 || d->n_ref < 0 || d->n_sync < 0 || d->n_obs < 0 ||
        (d->n_sweep && !d->sweep) || (d->n_ref && !d->ref) || (d->n_sync && !d->sync) || (d->n_obs && !d->obs))
        return fail(SCK_E_INVALID, "descriptor launch plans missing");
    if (d->rep_blocks < 1 || d->n_sweep % d->rep_blocks != 0)
        return fail(SCK_E_INVALID, "rep_blocks must be >= 1 and divide the sweep launches");
    const int obs_blocks = d->obs_blocks > 1 ? d->obs_blocks : 1;
    if (obs_blocks > 1 && (obs_blocks != d->rep_blocks || d->n_obs % obs_blocks != 0 || d->batch % obs_blocks != 0))
        return fail(SCK_E_INVALID, "obs_blocks must equal rep_blocks and divide the observation launches and the batch");
    std::string code;
    int rc = code_object(hip_src, cache_dir, code);
    if (rc) return rc;
    auto *h = new sck_handle();
    hipError_t e = hipModuleLoadData(&h->module, code.data());
    if (e != hipSuccess) {
        delete h;
        return fail(SCK_E_HIP, std::string("hipModuleLoadData: ") + hipGetErrorString(e));
    }
    for (int i = 0; i < d->n_kernels; ++i) {
        hipFunction_t f;
        e = hipModuleGetFunction(&f, h->module, d->kernel_names[i]);
        if (e != hipSuccess) {
            hipModuleUnload(h->module);
            delete h;
            return fail(SCK_E_HIP, std::string("kernel ") + d->kernel_names[i] + ": " + hipGetErrorString(e));
        }
        h->fn.push_back(f);
        h->names.push_back(d->kernel_names[i]);
        if (std::strcmp(d->kernel_names[i], "sck_sample") == 0) h->sample_kernel = i;
    }
    h->sweep.assign(d->sweep, d->sweep + d->n_sweep);
    h->ref.assign(d->ref, d->ref + d->n_ref);
    h->sync.assign(d->sync, d->sync + d->n_sync);
    h->obs.assign(d->obs, d->obs + d->n_obs);
    h->bump = *d->bump;
    if (d->resident) {
        h->has_resident = true;
        h->resident = *d->resident;
    }
    if (d->tiled) {
        if (d->mode != SCK_MODE_F32) {
            sck_destroy(h);
            return fail(SCK_E_INVALID, "the tiled sweep runs float32 spins");
        }
        h->has_tiled = true;
        h->tiled = *d->tiled;
        if (d->track_fold) {
            h->has_fold = true;
            h->fold = *d->track_fold;
        }
        if (h->tiled.kernel < 0 || h->tiled.kernel >= (int)h->fn.size()) {
            sck_destroy(h);
            return fail(SCK_E_INVALID, "tiled launch names an unknown kernel");
        }
        if (h->tiled.grid[1] == 0) {
            // persistent workgroups: as many per XCD column (grid.x) as are resident at once on the device
            int dev = 0, cus = 0, per = 0;
            hipError_t e1 = hipGetDevice(&dev);
            if (e1 == hipSuccess) e1 = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
            if (e1 == hipSuccess)
                e1 = hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per, h->fn[h->tiled.kernel],
                                                                       (int)h->tiled.block[0], h->tiled.shmem);
            if (e1 != hipSuccess) {
                sck_destroy(h);
                return fail(SCK_E_HIP, std::string("tiled occupancy: ") + hipGetErrorString(e1));
            }
            const unsigned cols = h->tiled.grid[0] ? h->tiled.grid[0] : 1;
            const unsigned wg = (unsigned)(cus > 0 ? cus : 1) * (unsigned)(per > 0 ? per : 1) / cols;
            h->tiled.grid[1] = wg > 0 ? wg : 1;
        }
        h->spin_bytes = (size_t)d->batch * d->n * d->b * d->L1 * d->L2 * d->L3 * sizeof(float);
    }
    h->mode = d->mode;
    h->n = d->n;
    h->b = d->b;
    h->batch = d->batch;
    h->words = d->words;
    h->obs_chunks = d->obs_chunks;
    h->rep_blocks = d->rep_blocks;
    h->obs_blocks = obs_blocks;
    *out = h;
    return SCK_OK;
}

int sck_destroy(sck_handle *h) {
    if (!h) return SCK_OK;
    if (h->module) hipModuleUnload(h->module);
    delete h;
    return SCK_OK;
}

int sck_sweep(sck_handle *h, const sck_args *args, int32_t n, int32_t class_lo, int32_t class_hi, void *stream) {
    if (!h || !args || n < 0) return fail(SCK_E_INVALID, "invalid sweep arguments");
    return do_sweeps(h, args, n, class_lo, class_hi, (hipStream_t)stream);
}

int sck_sweep_timed(sck_handle *h, const sck_args *args, int32_t n, void *stream, float *kernel_ms,
                    int32_t *launches) {
    if (!h || !args || n <= 0 || !kernel_ms || !launches) return fail(SCK_E_INVALID, "invalid timed-sweep arguments");
    hipStream_t s = (hipStream_t)stream;
    launch_timer t;
    int rc = do_sweeps(h, args, n, 0, -1, s, &t);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(s));
    float total = 0.0f;
    for (size_t i = 0; i + 1 < t.ev.size(); i += 2) {
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, t.ev[i], t.ev[i + 1]));
        total += ms;
    }
    *kernel_ms = total;
    *launches = (int32_t)(t.ev.size() / 2);
    return SCK_OK;
}

int sck_reference_sweep(sck_handle *h, const sck_args *args, int32_t n, void *stream) {
    if (!h || !args || n < 0) return fail(SCK_E_INVALID, "invalid reference-sweep arguments");
    hipStream_t s = (hipStream_t)stream;
    sck_args a = *args;
    for (int32_t i = 0; i < n; ++i) {
        a.sweep_offset = i;
        for (auto &l : h->ref) {
            int rc = launch(h, l, a, s);
            if (rc) return rc;
        }
    }
    return n > 0 ? bump(h, *args, n, s) : SCK_OK;
}

int sck_sync_sweep(sck_handle *h, const sck_args *args, int32_t n, int64_t spin_bytes, void *stream) {
    if (!h || !args || n < 0 || !args->dst || spin_bytes <= 0) return fail(SCK_E_INVALID, "invalid sync arguments");
    hipStream_t s = (hipStream_t)stream;
    sck_args a = *args;
    for (int32_t i = 0; i < n; ++i) {
        a.sweep_offset = i;
        HIP_TRY(hipMemcpyAsync(a.dst, a.spins, (size_t)spin_bytes, hipMemcpyDeviceToDevice, s));
        for (auto &l : h->sync) {
            int rc = launch(h, l, a, s);
            if (rc) return rc;
        }
        HIP_TRY(hipMemcpyAsync(a.spins, a.dst, (size_t)spin_bytes, hipMemcpyDeviceToDevice, s));
    }
    return n > 0 ? bump(h, *args, n, s) : SCK_OK;
}

int sck_observe(sck_handle *h, const sck_args *args, void *stream) {
    if (!h || !args || !args->obs || !args->obs_scratch) return fail(SCK_E_INVALID, "observe needs obs buffers");
    return do_observe(h, args, (hipStream_t)stream);
}

int sck_observe_noalloc(sck_handle *h, const sck_args *args, void *stream) {
    if (!h || !args || !args->obs) return fail(SCK_E_INVALID, "observe needs the obs output buffer");
    const size_t rows = (size_t)(h->batch / h->obs_blocks);       // every observation group reuses these rows
    const size_t need = rows * (size_t)h->obs_chunks * (size_t)(1 + h->b * h->n) * sizeof(double);
    if (need > SCK_OBS_SCRATCH_BYTES)
        return fail(SCK_E_INVALID, "observe_noalloc: the partial sums need " + std::to_string(need) +
                                       " bytes, more than the library's static scratch of " +
                                       std::to_string((size_t)SCK_OBS_SCRATCH_BYTES) + " bytes; use sck_observe");
    double *scratch = nullptr;
    int rc = obs_scratch_ptr(&scratch);
    if (rc) return rc;
    sck_args a = *args;
    a.obs_scratch = scratch;
    return do_observe(h, &a, (hipStream_t)stream);
}

int sck_graph_capture(sck_handle *h, const sck_args *args, int32_t k, int32_t observe, sck_graph **out) {
    if (!h || !args || k < 0 || !out) return fail(SCK_E_INVALID, "invalid graph arguments");
    auto *g = new sck_graph();
    hipError_t e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete g;
        return fail(SCK_E_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    e = hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        hipStreamDestroy(g->stream);
        delete g;
        return fail(SCK_E_HIP, std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
    }
    int rc;
    if ((observe & SCK_CAPTURE_OBSERVE) && h->obs_blocks > 1) {
        rc = do_sweeps_observed(h, args, k, g->stream);
    } else {
        // SCK_CAPTURE_IN_DST: whole tiled sweeps leave an odd count's result in args->dst (no copy back) and
        // the observation reads it there
        const bool keep = (observe & SCK_CAPTURE_IN_DST) != 0;
        rc = do_sweeps(h, args, k, 0, -1, g->stream, nullptr, !keep);
        if (!rc && (observe & SCK_CAPTURE_OBSERVE)) {
            sck_args o = *args;
            if (keep && (k & 1) && tiled_sweeps(h, args, k, 0, -1)) o.spins = args->dst;
            rc = do_observe(h, &o, g->stream);
        }
    }
    hipGraph_t graph = nullptr;
    hipError_t e2 = hipStreamEndCapture(g->stream, &graph);
    if (rc || e2 != hipSuccess) {
        if (graph) hipGraphDestroy(graph);
        hipStreamDestroy(g->stream);
        delete g;
        return rc ? rc : fail(SCK_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e2));
    }
    g->graph = graph;
    e = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        hipGraphDestroy(graph);
        hipStreamDestroy(g->stream);
        delete g;
        return fail(SCK_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    *out = g;
    return SCK_OK;
}

int sck_graph_replay(sck_graph *g, void *stream) {
    if (!g) return fail(SCK_E_INVALID, "null graph");
    HIP_TRY(hipGraphLaunch(g->exec, (hipStream_t)stream));
    return SCK_OK;
}

int sck_graph_destroy(sck_graph *g) {
    if (!g) return SCK_OK;
    if (g->exec) hipGraphExecDestroy(g->exec);
    if (g->graph) hipGraphDestroy(g->graph);
    if (g->stream) hipStreamDestroy(g->stream);
    delete g;
    return SCK_OK;
}

// ----------------------------------------------------------------------------- learned proposal width
static const int kMlpH = 32;
static const int kMlpParams = kMlpH * 3 + kMlpH + kMlpH * kMlpH + kMlpH + kMlpH + 1;
static const int kMlpChunk = 256;

int sck_mlp_nparam(void) { return kMlpParams; }

int64_t sck_mlp_workspace_bytes(int32_t nsample, int32_t ndraw) {
    (void)ndraw;
    int64_t nrec = (int64_t)nsample;
    int64_t nchunk = (nrec + kMlpChunk - 1) / kMlpChunk;
    return 4 * (nrec * (5 + 4 * kMlpH) + nchunk * (kMlpParams + 1));
}

int sck_mlp_train_step(float *params, float *adam_m, float *adam_v, const sck_mlp_sample *samples, int32_t nsample,
                       const sck_mlp_config *cfg, const sck_mlp_adam_config *adam, void *workspace, float *grad_out,
                       void *stream) {
    if (!params || !samples || nsample <= 0 || !cfg || cfg->ndraw <= 0 || !adam || !workspace || !grad_out)
        return fail(SCK_E_INVALID, "invalid training-step arguments");
    if (adam->lr != 0.0f && (!adam_m || !adam_v || adam->step < 1))
        return fail(SCK_E_INVALID, "an Adam step needs its moments and a 1-based step");
    hipStream_t s = (hipStream_t)stream;
    int32_t nrec = nsample;
    float *rec = (float *)workspace;
    int32_t nchunk = (nrec + kMlpChunk - 1) / kMlpChunk;
    float *partial = rec + (int64_t)nrec * (5 + 4 * kMlpH);
    sck_mlp_config c = *cfg;
    hipLaunchKernelGGL(sck_mlp_fwdbwd_k, dim3((nrec + 127) / 128), dim3(128), 0, s, params, samples, nsample, c, rec);
    HIP_TRY(hipGetLastError());
    int32_t chunk = kMlpChunk;
    hipLaunchKernelGGL(sck_mlp_reduce_k, dim3(nchunk, (kMlpParams + 1 + 1023) / 1024), dim3(1024), 0, s, rec, nrec,
                       chunk, partial);
    HIP_TRY(hipGetLastError());
    sck_mlp_adam_config ac = *adam;
    hipLaunchKernelGGL(sck_mlp_adam_k, dim3((kMlpParams + 1 + 255) / 256), dim3(256), 0, s, params, adam_m, adam_v,
                       grad_out, partial, nchunk, ac);
    HIP_TRY(hipGetLastError());
    return SCK_OK;
}

int sck_mlp_table(const float *params, const sck_mlp_config *cfg, const sck_width_grid *grid, float *table,
                  void *stream) {
    if (!params || !cfg || !grid || !table) return fail(SCK_E_INVALID, "invalid table arguments");
    int total = grid->n[0] * grid->n[1] * grid->n[2];
    if (total <= 0) return fail(SCK_E_INVALID, "empty width grid");
    hipLaunchKernelGGL(sck_mlp_table_k, dim3((total + 127) / 128), dim3(128), 0, (hipStream_t)stream, params, *cfg,
                       *grid, table);
    HIP_TRY(hipGetLastError());
    return SCK_OK;
}

int sck_sample_sites(sck_handle *h, const sck_args *args, const int32_t *rep_site, int32_t nsample,
                     sck_mlp_sample *out, void *stream) {
    if (!h || !args || !rep_site || nsample <= 0 || !out) return fail(SCK_E_INVALID, "invalid sampling arguments");
    if (h->sample_kernel < 0) return fail(SCK_E_INVALID, "this module has no sampling kernel (vector models only)");
    sck_args a = *args;
    const int32_t *rs = rep_site;
    int32_t n = nsample;
    sck_mlp_sample *o = out;
    void *params[] = {&a, &rs, &n, &o};
    HIP_TRY(hipModuleLaunchKernel(h->fn[h->sample_kernel], (nsample + 127) / 128, 1, 1, 128, 1, 1, 0,
                                  (hipStream_t)stream, params, nullptr));
    return SCK_OK;
}

int sck_time_replays(sck_graph *g, int32_t reps, void *stream, float *ms) {
    if (!g || reps <= 0 || !ms) return fail(SCK_E_INVALID, "invalid timing arguments");
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t a = nullptr, b = nullptr;
    hipError_t e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    if (e == hipSuccess) e = hipEventRecord(a, s);
    for (int32_t i = 0; i < reps && e == hipSuccess; ++i) e = hipGraphLaunch(g->exec, s);
    if (e == hipSuccess) e = hipEventRecord(b, s);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, a, b);
    if (a) hipEventDestroy(a);
    if (b) hipEventDestroy(b);
    if (e != hipSuccess) return fail(SCK_E_HIP, std::string("replay timing: ") + hipGetErrorString(e));
    return SCK_OK;
}

}  // extern "C"
