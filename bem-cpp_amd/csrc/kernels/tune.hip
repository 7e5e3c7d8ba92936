// tune.hip -- autotuner for the GEMM tile shapes of the factorisation and the
// solves.  The reference leaves these choices to cuSOLVER; here the LU
// trailing updates (m x m x 256 products) and the GEMM-based triangular solves
// and refinement residuals (n x nrhs products) run on the configurations of
// kernels/gemm.h, and this routine times every configuration at the shapes of
// the context's system and keeps the fastest per shape class.
//
// Every configuration accumulates K in the same order (slices of 16, four
// MFMA steps of 4), so the choice changes speed only, not results.
#include <hip/hip_runtime.h>

#include <sys/stat.h>
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../bem_internal.h"
#include "gemm.h"
#include "gemm_nt.h"

namespace {

template <typename T>
__global__ void k_fill_tune(T* p, long n, unsigned seed) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned x = (unsigned)i * 2654435761u ^ seed;
    x ^= x >> 13;
    x *= 0x5bd1e995u;
    x ^= x >> 15;
    p[i] = (T)((x & 0xffffff) * (1.0 / 16777216.0) - 0.5);
}

struct Shape {
    int M, N, K;
};

// Best configuration and its rate (TFLOP/s) for one shape; rates[cfg] filled.
// With a plan, every configuration runs under its headroom grid cap.
template <typename T>
int time_configs(hipStream_t s, const Shape& sh, T* buf, double* rates, const GemmPlan* capped) {
    T* Ar = buf;
    T* Ai = Ar + (size_t)sh.M * sh.K;
    T* Br = Ai + (size_t)sh.M * sh.K;
    T* Bi = Br + (size_t)sh.K * sh.N;
    T* Cr = Bi + (size_t)sh.K * sh.N;
    T* Ci = Cr + (size_t)sh.M * sh.N;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return -1;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return -1;
    }
    const double flops = (sizeof(T) == 4 ? 6.0 : 8.0) * sh.M * (double)sh.N * sh.K;
    int best = 0;
    for (int cfg = 0; cfg < kGemmCfgs; cfg++) {
        auto run = [&]() {
            cgemm_any<T>(cfg, s, sh.M, sh.N, sh.K, Ar, Ai, sh.K, Br, Bi, sh.N, Cr, Ci, sh.N, 0, 1, 0, 0, 0,
                         capped ? gemm_headroom_cap(*capped, cfg) : 0);
        };
        run();  // warm-up (code load, caches)
        const int reps = 3;
        (void)hipEventRecord(e0, s);
        for (int r = 0; r < reps; r++) run();
        (void)hipEventRecord(e1, s);
        (void)hipEventSynchronize(e1);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        rates[cfg] = ms > 0.0f ? flops * reps / (ms * 1e-3) / 1e12 : 0.0;
        if (rates[cfg] > rates[best]) best = cfg;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return best;
}

// Results shared by the contexts of one process: (device, precision, npad, ldr).
std::mutex g_mu;
std::map<std::tuple<int, int, int, int>, GemmPlan> g_cache;

// ---------------------------------------------------------------------------
// Plans persisted across processes: a JSON file (BEM_TUNE_CACHE, else
// $XDG_CACHE_HOME/bem_amd/gemm_plans.json, else ~/.cache/bem_amd/..., else,
// without a writable home, /tmp/bem_amd-<uid>/bem_amd/...; an empty
// BEM_TUNE_CACHE turns the file off), one plan per line, keyed by the device
// (architecture name and CU count), the precision, npad and ldr.  A plan found
// there replaces the timing of the tiles and, when it was chosen by trial
// factorisations, the trials too.  kPlanFormat changes whenever the tile
// configurations change, so plans of another build are ignored.
// ---------------------------------------------------------------------------
constexpr int kPlanFormat = 5;

std::string plan_file() {
    if (const char* e = std::getenv("BEM_TUNE_CACHE")) return e;
    std::string base;
    if (const char* x = std::getenv("XDG_CACHE_HOME"); x && *x && ::access(x, W_OK) == 0)
        base = x;
    else if (const char* h = std::getenv("HOME"); h && *h && ::access(h, W_OK) == 0)
        base = std::string(h) + "/.cache";
    else  // no writable home (service accounts): a per-user folder in /tmp
        base = "/tmp/bem_amd-" + std::to_string((long)::getuid());
    return base + "/bem_amd/gemm_plans.json";
}

std::string device_tag(int device) {
    hipDeviceProp_t p{};
    if (hipGetDeviceProperties(&p, device) != hipSuccess) return "unknown";
    return std::string(p.gcnArchName) + "/" + std::to_string(p.multiProcessorCount);
}

std::string plan_line(const std::string& dev, int prec, int npad, int ldr, const GemmPlan& p) {
    char b[1024];
    int k = std::snprintf(b, sizeof b,
                          "{\"format\": %d, \"device\": \"%s\", \"prec\": %d, \"npad\": %d, \"ldr\": %d, "
                          "\"large\": %d, \"skinny\": %d, \"panel\": %d, \"nt\": %d, \"fill\": %ld, "
                          "\"trialed\": %d, \"trial_ms\": [",
                          kPlanFormat, dev.c_str(), prec, npad, ldr, p.large, p.skinny, p.panel, p.nt, p.fill,
                          p.trialed ? 1 : 0);
    for (int i = 0; i < GemmPlan::kMaxCfgs; i++)
        k += std::snprintf(b + k, sizeof b - k, "%s%.17g", i ? ", " : "", p.trial_ms[i]);
    k += std::snprintf(b + k, sizeof b - k, "], \"alone_tflops\": [");
    for (int i = 0; i < GemmPlan::kMaxCfgs; i++)
        k += std::snprintf(b + k, sizeof b - k, "%s%.17g", i ? ", " : "", p.alone_tflops[i]);
    std::snprintf(b + k, sizeof b - k, "]}");
    return b;
}

// Number after "key": in line (false when absent).
bool json_num(const std::string& line, const char* key, double& v) {
    const std::string k = std::string("\"") + key + "\":";
    const size_t at = line.find(k);
    if (at == std::string::npos) return false;
    char* end = nullptr;
    v = std::strtod(line.c_str() + at + k.size(), &end);
    return end != line.c_str() + at + k.size();
}
bool json_array(const std::string& line, const char* key, double* v, int n) {
    const std::string k = std::string("\"") + key + "\": [";
    size_t at = line.find(k);
    if (at == std::string::npos) return false;
    const char* q = line.c_str() + at + k.size();
    for (int i = 0; i < n; i++) {
        char* end = nullptr;
        v[i] = std::strtod(q, &end);
        if (end == q) return false;
        q = end;
        while (*q == ',' || *q == ' ') q++;
    }
    return true;
}
bool json_str(const std::string& line, const char* key, std::string& v) {
    const std::string k = std::string("\"") + key + "\": \"";
    const size_t at = line.find(k);
    if (at == std::string::npos) return false;
    const size_t e = line.find('"', at + k.size());
    if (e == std::string::npos) return false;
    v = line.substr(at + k.size(), e - (at + k.size()));
    return true;
}

std::vector<std::string> plan_file_lines(const std::string& path) {
    std::vector<std::string> out;
    std::ifstream f(path);
    std::string line;
    while (std::getline(f, line))
        if (line.rfind("{\"format\"", 0) == 0) out.push_back(line);
    return out;
}

// The persisted plan of (device, prec, npad, ldr) into p; false if none.
bool plan_load(int device, int prec, int npad, int ldr, GemmPlan& p) {
    const std::string path = plan_file();
    if (path.empty()) return false;
    const std::string dev = device_tag(device);
    for (const std::string& line : plan_file_lines(path)) {
        double fmt = 0, pr = 0, np = 0, lr = 0, v = 0;
        std::string d;
        if (!json_num(line, "format", fmt) || (int)fmt != kPlanFormat || !json_str(line, "device", d) || d != dev ||
            !json_num(line, "prec", pr) || (int)pr != prec || !json_num(line, "npad", np) || (int)np != npad ||
            !json_num(line, "ldr", lr) || (int)lr != ldr)
            continue;
        GemmPlan q = p;
        if (!json_num(line, "large", v)) continue;
        q.large = (int)v;
        if (!json_num(line, "skinny", v)) continue;
        q.skinny = (int)v;
        if (!json_num(line, "panel", v)) continue;
        q.panel = (int)v;
        if (!json_num(line, "nt", v)) continue;
        q.nt = (int)v;
        if (!json_num(line, "fill", v)) continue;
        q.fill = (long)v;
        if (!json_num(line, "trialed", v)) continue;
        q.trialed = v != 0.0;
        if (!json_array(line, "trial_ms", q.trial_ms, GemmPlan::kMaxCfgs)) continue;
        if (!json_array(line, "alone_tflops", q.alone_tflops, GemmPlan::kMaxCfgs)) continue;
        if (q.large < 0 || q.large >= kGemmCfgs || q.skinny < 0 || q.skinny >= kGemmCfgs || q.panel < 0 ||
            q.panel >= kGemmCfgs || q.nt < 0 || q.nt >= kNtCfgs)
            continue;
        q.tuned = true;
        q.npad = npad;
        q.cap = 0;
        p = q;
        return true;
    }
    return false;
}

// Writes (or replaces) the plan's line: the file is rewritten through a
// temporary file and a rename, so a reader never sees half a file.
void plan_save(int device, int prec, int npad, int ldr, const GemmPlan& p) {
    const std::string path = plan_file();
    if (path.empty()) return;
    const std::string dev = device_tag(device);
    std::vector<std::string> keep;
    for (const std::string& line : plan_file_lines(path)) {
        double fmt = 0, pr = 0, np = 0, lr = 0;
        std::string d;
        const bool same = json_num(line, "format", fmt) && (int)fmt == kPlanFormat && json_str(line, "device", d) &&
                          d == dev && json_num(line, "prec", pr) && (int)pr == prec && json_num(line, "npad", np) &&
                          (int)np == npad && json_num(line, "ldr", lr) && (int)lr == ldr;
        if (!same) keep.push_back(line);
    }
    keep.push_back(plan_line(dev, prec, npad, ldr, p));
    const size_t slash = path.rfind('/');
    if (slash != std::string::npos) {  // create the directory chain
        const std::string dir = path.substr(0, slash);
        for (size_t i = 1; i <= dir.size(); i++)
            if (i == dir.size() || dir[i] == '/') (void)::mkdir(dir.substr(0, i).c_str(), 0755);
    }
    const std::string tmp = path + ".tmp" + std::to_string((long)::getpid());
    {
        std::ofstream f(tmp);
        if (!f) return;
        f << "{\"bem_amd_gemm_plans\": " << kPlanFormat << ", \"plans\": [\n";
        for (size_t i = 0; i < keep.size(); i++) f << keep[i] << (i + 1 < keep.size() ? ",\n" : "\n");
        f << "]}\n";
        if (!f) {
            f.close();
            (void)std::remove(tmp.c_str());
            return;
        }
    }
    if (std::rename(tmp.c_str(), path.c_str()) != 0) (void)std::remove(tmp.c_str());
}

template <typename T>
int tune_precision(bem_ctx* c, int npad, int ldr, double* rates_large, double* rates_skinny) {
    GemmPlan& plan = c->gemm_plan[sizeof(T) == 4 ? 0 : 1];
    const auto key = std::make_tuple(c->device, (int)sizeof(T), npad, ldr);
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_cache.find(key);
        if (it != g_cache.end() && !rates_large) {
            plan = it->second;  // carries npad
            return 0;
        }
        if (!rates_large && plan_load(c->device, (int)sizeof(T), npad, ldr, plan)) {
            gemm_plan_occupancy(c, sizeof(T) == 4);
            g_cache[key] = plan;
            return 0;
        }
    }
    // trailing update (capped so the timing stays short) and GEMM-solve shapes
    const Shape large{std::min(npad, 8192), std::min(npad, 8192), 256};
    const Shape skinny{npad, ldr, 256};
    size_t elems = 0;
    for (const Shape& sh : {large, skinny})
        elems = std::max(elems, 2 * ((size_t)sh.M * sh.K + (size_t)sh.K * sh.N + (size_t)sh.M * sh.N));
    T* buf = nullptr;
    if (hipMalloc(&buf, elems * sizeof(T)) != hipSuccess) return bem_fail(c, BEM_ERR_NOMEM, "autotune buffers");
    hipLaunchKernelGGL(k_fill_tune<T>, dim3((unsigned)((elems + 255) / 256)), dim3(256), 0, c->stream, buf,
                       (long)elems, 7u);
    gemm_plan_occupancy(c, sizeof(T) == 4);
    double rl[kGemmCfgs] = {0}, rs[kGemmCfgs] = {0};
    const int bl = time_configs<T>(c->stream, large, buf, rl, nullptr);
    const int bs = time_configs<T>(c->stream, skinny, buf, rs, nullptr);
    const hipError_t e = hipStreamSynchronize(c->stream);
    (void)hipFree(buf);
    if (e != hipSuccess || bl < 0 || bs < 0)
        return bem_fail(c, BEM_ERR_HIP, std::string("autotune: ") + hipGetErrorString(e));
    plan.large = bl;
    plan.skinny = bs;
    for (int i = 0; i < kGemmCfgs; i++) plan.alone_tflops[i] = rl[i];
    plan.cap = 0;
    plan.trialed = false;  // the first factorisation of this size times whole trials
    for (double& t : plan.trial_ms) t = 0.0;
    plan.fill = 2L * plan.cus;  // fewer blocks of the large tile than two per CU: use the skinny tile
    plan.tuned = true;
    plan.npad = npad;
    if (rates_large)
        for (int i = 0; i < kGemmCfgs; i++) {
            rates_large[i] = rl[i];
            rates_skinny[i] = rs[i];
        }
    std::lock_guard<std::mutex> lk(g_mu);
    g_cache[key] = plan;
    plan_save(c->device, (int)sizeof(T), npad, ldr, plan);
    return 0;
}

}  // namespace

int gemm_autotune_launch(bem_ctx* c, int npad, int nrhs, double* out, int nout) {
    const int ldr = round_up(std::max(nrhs, 1), 16);
    double r[4][kGemmCfgs] = {};
    c->gemm_plan[0].cap = c->gemm_plan[1].cap = 0;
    const bool want = out && nout > 0;
    if (int rc = tune_precision<float>(c, npad, ldr, want ? r[0] : nullptr, want ? r[1] : nullptr)) return rc;
    if (int rc = tune_precision<double>(c, npad, ldr, want ? r[2] : nullptr, want ? r[3] : nullptr)) return rc;
    if (want) {
        // [precision][shape][cfg] rates, then the chosen (large, skinny) per precision
        int k = 0;
        for (int a = 0; a < 4; a++)
            for (int i = 0; i < kGemmCfgs && k < nout; i++) out[k++] = r[a][i];
        for (int p = 0; p < 2; p++) {
            if (k < nout) out[k++] = c->gemm_plan[p].large;
            if (k < nout) out[k++] = c->gemm_plan[p].skinny;
        }
        return k;
    }
    return 0;
}

void gemm_plan_occupancy(bem_ctx* c, bool f32) {
    GemmPlan& plan = c->gemm_plan[f32 ? 0 : 1];
    if (plan.occ_known) return;
    (void)hipDeviceGetAttribute(&plan.cus, hipDeviceAttributeMultiprocessorCount, c->device);
    for (int i = 0; i < kGemmCfgs; i++) plan.occ[i] = f32 ? cgemm_occupancy<float>(i) : cgemm_occupancy<double>(i);
    plan.occ_known = true;
}

void gemm_cache_store(bem_ctx* c, bool f32) {
    const GemmPlan& plan = c->gemm_plan[f32 ? 0 : 1];
    const int ldr = round_up(std::max(c->ldr > 0 ? c->ldr : 128, 1), 16);
    std::lock_guard<std::mutex> lk(g_mu);
    g_cache[std::make_tuple(c->device, f32 ? 4 : 8, plan.npad, ldr)] = plan;
    plan_save(c->device, f32 ? 4 : 8, plan.npad, ldr, plan);
}

// Fixes the plan of one precision (0 FP32, 1 FP64) instead of tuning it.
int gemm_set_plan(bem_ctx* c, int prec, int large, int skinny, long cap) {
    if (prec < 0 || prec > 1 || large < 0 || large >= kGemmCfgs || skinny < 0 || skinny >= kGemmCfgs || cap < 0)
        return bem_fail(c, BEM_ERR_ARG, "gemm_plan: bad configuration");
    if (c->npad <= 0) return bem_fail(c, BEM_ERR_STATE, "gemm_plan: no system size yet (set a mesh or a matrix)");
    GemmPlan& plan = c->gemm_plan[prec];
    gemm_plan_occupancy(c, prec == 0);
    plan.large = large;
    plan.skinny = skinny;
    plan.cap = cap;
    plan.tuned = true;
    plan.trialed = true;  // fixed: no factorisation trials
    for (double& t : plan.trial_ms) t = 0.0;
    plan.npad = c->npad;
    return 0;
}

// Tunes the precision about to be factored once per (device, size) unless the
// environment sets BEM_AUTOTUNE=0 (then the defaults of GemmPlan stay).
int gemm_ensure_tuned(bem_ctx* c, bool f32) {
    GemmPlan& plan = c->gemm_plan[f32 ? 0 : 1];
    if (plan.tuned && plan.npad == c->npad) return 0;
    const char* env = std::getenv("BEM_AUTOTUNE");
    if (env && env[0] == '0') {
        gemm_plan_occupancy(c, f32);
        plan.tuned = true;
        plan.trialed = true;
        plan.npad = c->npad;
        return 0;
    }
    const int ldr = round_up(std::max(c->ldr > 0 ? c->ldr : 128, 1), 16);
    return f32 ? tune_precision<float>(c, c->npad, ldr, nullptr, nullptr)
               : tune_precision<double>(c, c->npad, ldr, nullptr, nullptr);
}

namespace {
__global__ __launch_bounds__(256) void k_copy16(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}
}  // namespace

int device_copy_launch(bem_ctx* c, size_t n16) {
    float4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, n16 * 16) != hipSuccess) return bem_fail(c, BEM_ERR_NOMEM, "device_copy: source");
    if (hipMalloc(&b, n16 * 16) != hipSuccess) {
        (void)hipFree(a);
        return bem_fail(c, BEM_ERR_NOMEM, "device_copy: destination");
    }
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device);
    hipError_t e = hipMemsetAsync(a, 0, n16 * 16, c->stream);
    if (!e) {
        hipLaunchKernelGGL(k_copy16, dim3(8 * cus), dim3(256), 0, c->stream, a, b, n16);
        e = hipStreamSynchronize(c->stream);
    }
    (void)hipFree(a);
    (void)hipFree(b);
    if (e != hipSuccess) return bem_fail(c, BEM_ERR_HIP, std::string("device_copy: ") + hipGetErrorString(e));
    return 0;
}
