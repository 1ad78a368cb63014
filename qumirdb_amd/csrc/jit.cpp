// Kernels specialized at run time. An operator whose layout is fixed when it
// is compiled (the aggregate's keys, state words, predicate and computed
// arguments) gets a kernel generated for it, as the reference's compiler
// generates one per query (qdb/kernel/compiler.cpp): the generated source
// includes the library's own device headers, which are built into the
// library, and is compiled with hipRTC for the device's architecture. Each
// distinct source is compiled once per process.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <poll.h>
#include <signal.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/wait.h>
#include <unistd.h>

#include <cerrno>
#include <algorithm>
#include <chrono>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "jit.h"
#include "runtime.h"

namespace qdbgpu {

namespace {

// The device headers, embedded by the build (jit_headers.inc defines
// kJitHeaderNames / kJitHeaderTexts / kJitHeaderCount).
#include "jit_headers.inc"

struct JitEntry {
    hipModule_t module = nullptr;
    std::string error;
    double compile_ms = 0.0;
    std::unordered_map<std::string, hipFunction_t> fns;
};

struct JitCache {
    std::mutex mu;
    std::unordered_map<std::string, JitEntry> map;
    int compiles = 0;
    double total_ms = 0.0;
};

JitCache& cache()
{
    static JitCache* c = new JitCache();
    return *c;
}

std::string device_arch()
{
    hipDeviceProp_t p;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess) {
        (void)hipGetLastError();
        return "gfx950";
    }
    std::string a = p.gcnArchName;
    const size_t colon = a.find(':');
    return colon == std::string::npos ? a : a.substr(0, colon);
}

// Compiles `src` to a code object in this process; false with the log in
// *err.
bool compile_here(const std::string& src, const std::string& arch, std::vector<char>* code, std::string* err)
{
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "qdb_specialized.hip", kJitHeaderCount, kJitHeaderTexts, kJitHeaderNames) !=
        HIPRTC_SUCCESS) {
        *err = "hiprtcCreateProgram failed";
        return false;
    }
    const std::string arch_opt = "--offload-arch=" + arch;
    // The packed-FP32 feature stays off in run-time kernels too.
    const char* opts[] = {arch_opt.c_str(), "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
    const hiprtcResult r = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
    size_t log_size = 0;
    (void)hiprtcGetProgramLogSize(prog, &log_size);
    std::string log(log_size, '\0');
    if (log_size) (void)hiprtcGetProgramLog(prog, &log[0]);
    if (r != HIPRTC_SUCCESS) {
        *err = std::string("hipRTC: ") + hiprtcGetErrorString(r) + "\n" + log;
        (void)hiprtcDestroyProgram(&prog);
        return false;
    }
    size_t size = 0;
    (void)hiprtcGetCodeSize(prog, &size);
    code->resize(size);
    (void)hiprtcGetCode(prog, code->data());
    (void)hiprtcDestroyProgram(&prog);
    return size > 0;
}

bool write_all(int fd, const void* p, size_t n)
{
    const char* c = (const char*)p;
    while (n) {
        const ssize_t k = write(fd, c, n);
        if (k <= 0) return false;
        c += k;
        n -= (size_t)k;
    }
    return true;
}

// Reads n bytes before `deadline`; false on end of file, error or time out
// (*timed_out set).
bool read_all(int fd, void* p, size_t n, std::chrono::steady_clock::time_point deadline, bool* timed_out)
{
    char* c = (char*)p;
    while (n) {
        const auto left = std::chrono::duration_cast<std::chrono::milliseconds>(deadline - std::chrono::steady_clock::now());
        if (left.count() <= 0) {
            *timed_out = true;
            return false;
        }
        struct pollfd pfd = {fd, POLLIN, 0};
        const int r = poll(&pfd, 1, (int)std::min<long long>(left.count(), 1000));
        if (r < 0 && errno == EINTR) continue;
        if (r < 0) return false;
        if (r == 0) continue;
        const ssize_t k = read(fd, c, n);
        if (k < 0 && errno == EINTR) continue;
        if (k <= 0) return false;
        c += k;
        n -= (size_t)k;
    }
    return true;
}

// Seconds a run-time compile may take (QDB_GPU_JIT_TIMEOUT_S, default 120);
// past it the compiler process is killed and the generic kernels run.
int jit_timeout_s()
{
    const char* e = getenv("QDB_GPU_JIT_TIMEOUT_S");
    const int v = e ? atoi(e) : 0;
    return v > 0 ? v : 120;
}

// Compiles `src` in a forked child that runs only the compiler and never
// touches the GPU: a fault inside the compiler then ends the child, and
// the caller gets a failed compile (and runs the generic kernels) instead
// of a crash. QDB_GPU_JIT_INPROCESS=1 compiles in this process.
bool compile_source(const std::string& src, const std::string& arch, std::vector<char>* code, std::string* err)
{
    const char* inproc = getenv("QDB_GPU_JIT_INPROCESS");
    if (inproc && inproc[0] == '1') return compile_here(src, arch, code, err);
    int fds[2];
    if (pipe(fds) != 0) return compile_here(src, arch, code, err);
    fflush(nullptr);
    const pid_t pid = fork();
    if (pid < 0) {
        close(fds[0]);
        close(fds[1]);
        return compile_here(src, arch, code, err);
    }
    if (pid == 0) {
        close(fds[0]);
        std::vector<char> c;
        std::string e;
        const uint64_t ok = compile_here(src, arch, &c, &e) ? 1 : 0;
        const uint64_t hdr[3] = {ok, (uint64_t)c.size(), (uint64_t)e.size()};
        const bool sent = write_all(fds[1], hdr, sizeof(hdr)) && write_all(fds[1], c.data(), c.size()) &&
            write_all(fds[1], e.data(), e.size());
        close(fds[1]);
        _exit(sent ? 0 : 3);
    }
    close(fds[1]);
    const auto deadline = std::chrono::steady_clock::now() + std::chrono::seconds(jit_timeout_s());
    bool timed_out = false;
    uint64_t hdr[3] = {0, 0, 0};
    bool got = read_all(fds[0], hdr, sizeof(hdr), deadline, &timed_out);
    if (got) {
        code->resize(hdr[1]);
        std::string e(hdr[2], '\0');
        got = read_all(fds[0], code->data(), hdr[1], deadline, &timed_out) && read_all(fds[0], &e[0], hdr[2], deadline, &timed_out);
        *err = e;
    }
    close(fds[0]);
    // A child that neither answered nor exited in time (a lock held by
    // another thread when the process forked, a compiler that does not end)
    // is killed, so the caller never waits on it for good.
    if (timed_out) kill(pid, SIGKILL);
    int status = 0;
    while (waitpid(pid, &status, 0) < 0 && errno == EINTR) {
    }
    if (!got || !WIFEXITED(status) || WEXITSTATUS(status) != 0) {
        char b[96];
        if (timed_out) snprintf(b, sizeof(b), "the run-time compiler took longer than %d s", jit_timeout_s());
        else if (WIFSIGNALED(status)) snprintf(b, sizeof(b), "the run-time compiler ended with signal %d", WTERMSIG(status));
        else snprintf(b, sizeof(b), "the run-time compiler process failed (status %d)", status);
        *err = b;
        code->clear();
        return false;
    }
    return hdr[0] == 1 && !code->empty();
}

}  // namespace

bool jit_enabled()
{
    const char* e = getenv("QDB_GPU_JIT");
    return !(e && e[0] == '0');
}

hipFunction_t jit_function(const std::string& src, const char* name, std::string* err)
{
    JitCache& c = cache();
    std::lock_guard<std::mutex> lock(c.mu);
    auto it = c.map.find(src);
    if (it == c.map.end()) {
        JitEntry e;
        const auto t0 = std::chrono::steady_clock::now();
        std::vector<char> code;
        if (compile_source(src, device_arch(), &code, &e.error)) {
            const hipError_t h = hipModuleLoadData(&e.module, code.data());
            if (h != hipSuccess) {
                e.error = std::string("module load: ") + hipGetErrorString(h);
                e.module = nullptr;
                (void)hipGetLastError();
            }
        }
        e.compile_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        c.compiles += 1;
        c.total_ms += e.compile_ms;
        const char* log = getenv("QDB_GPU_JIT_LOG");
        if (log && log[0] == '1') {
            fprintf(stderr, "[qdb_gpu jit] compiled in %.0f ms%s%s\n", e.compile_ms, e.module ? "" : ", failed: ",
                e.module ? "" : e.error.c_str());
        }
        it = c.map.emplace(src, std::move(e)).first;
    }
    JitEntry& e = it->second;
    if (!e.module) {
        if (err) *err = e.error;
        return nullptr;
    }
    auto f = e.fns.find(name);
    if (f != e.fns.end()) return f->second;
    hipFunction_t fn = nullptr;
    const hipError_t h = hipModuleGetFunction(&fn, e.module, name);
    if (h != hipSuccess) {
        (void)hipGetLastError();
        if (err) *err = std::string("no kernel ") + name;
        fn = nullptr;
    }
    e.fns.emplace(name, fn);
    return fn;
}

bool jit_compile_only(const std::string& src, const std::string& arch, std::string* err)
{
    std::vector<char> code;
    return compile_source(src, arch, &code, err);
}

void jit_stats(int* compiles, double* total_ms)
{
    JitCache& c = cache();
    std::lock_guard<std::mutex> lock(c.mu);
    *compiles = c.compiles;
    *total_ms = c.total_ms;
}

}  // namespace qdbgpu

extern "C" int64_t qdb_gpu_jit_stats(double* total_ms)
{
    int n = 0;
    double ms = 0.0;
    qdbgpu::jit_stats(&n, &ms);
    if (total_ms) *total_ms = ms;
    return n;
}
