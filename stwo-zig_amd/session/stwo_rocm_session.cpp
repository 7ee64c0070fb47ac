// stwo_rocm_session: persistent JSONL prover session over the stwo_rocm C-ABI.
//
// Serving front door of the device backend, after the reference's sequential session loop
// (src/tools/metal_prover_session/app.zig:73-211) and its strict frame protocol
// (src/tools/metal_session/protocol.zig), with one report per request in the layout of the native
// runner's report schema 7 (src/prover/native/report.zig; statistics.zig for the summaries).  The
// device context (its stream, arena, twiddle tower and domain tables) lives for the whole session,
// so requests after the first reuse what earlier ones built.
//
// Frames, one JSON object per line on stdin, answered one per line on stdout:
//   {"protocol":"stwo-zig-rocm-prover-session","version":1,"type":"prove","sequence":S,
//    "request_id":ID,"workload":{"example":...},"proof_protocol":"smoke|functional|secure",
//    "warmups":W,"samples":K,"outputs":{"proof":ABS,"report":ABS},"budget_gib":"G"}
//   {"protocol":...,"version":1,"type":"shutdown","next_sequence":N}
// Answers: "ready" once the device is up, "result" per proof request, "closed" after shutdown,
// "error" for the first failure, after which the session ends (exit status 1): it fails closed
// like the reference.  Sequences start at 0 and increase by one; outputs must not exist and are
// published only whole (written aside, then hard-linked into place).
//
// With --verifier PATH [--verifier-arg A]..., every warmup and sample proof is checked by running
// PATH A... --statement STATEMENT.json --proof PROOF (exit status 0 = accepted) before anything is
// published, the reference's external-verifier step (app.zig:121-133).  --check answers each valid
// request with a "checked" frame (descriptor digest, admission) without a device.
#include <dlfcn.h>
#include <errno.h>
#include <fcntl.h>
#include <spawn.h>
#include <stdio.h>
#include <stdlib.h>
#include <sys/resource.h>
#include <sys/stat.h>
#include <sys/utsname.h>
#include <sys/wait.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <string>
#include <vector>

#include "../../include/stwo_rocm.h"
#include "json.hpp"
#include "sha256.hpp"

extern char** environ;

// Source revision of the build (set by the build script; "unknown" outside a git checkout).
#ifndef STWO_ROCM_GIT_COMMIT
#define STWO_ROCM_GIT_COMMIT "unknown"
#endif
#ifndef STWO_ROCM_GIT_DIRTY
#define STWO_ROCM_GIT_DIRTY 1
#endif

namespace {

const char* const kProtocol = "stwo-zig-rocm-prover-session";
const uint64_t kVersion = 1;
const size_t kMaxFrame = 4u << 20;  // protocol.zig:8
const uint32_t kMaxWarmups = 30;    // native/config.zig:207
const uint32_t kMaxSamples = 21;    // native/config.zig:415
const uint32_t kMinHeadlineWarmups = 10;  // native/config.zig:206
const char* const kTimingScope = "prove_call_start_to_hashed_verified_proof";

struct Fail {
    std::string code, msg;
};
[[noreturn]] void fail(const std::string& code, const std::string& msg) { throw Fail{code, msg}; }

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// ------------------------------------------------------------------ request model
struct Workload {
    std::string example;  // wide_fibonacci | poseidon | xor | xor_logup
    uint32_t log_rows = 0, width = 0;  // wide: log_n_rows, sequence_len; poseidon: log_n_instances;
                                       // xor: log_size, log_step; xor_logup: log_n
    uint64_t value = 0;                // xor: offset; xor_logup: seed
};

struct ProofProtocol {  // native/config.zig:140-160
    std::string name;
    uint32_t pow_bits = 10, log_blowup_factor = 1, log_last_layer_degree_bound = 0, n_queries = 3,
             fold_step = 1;
};

struct Request {
    uint64_t sequence = 0;
    std::string request_id;
    Workload w;
    ProofProtocol proto;
    uint32_t warmups = 0, samples = 1;
    std::string proof_out, report_out, budget_gib;
    uint64_t budget_bytes = 0;
};

struct Geometry {  // native/examples.zig:17-123
    uint32_t trace_log_rows;
    uint64_t trace_rows;
    uint32_t committed_trees;
    uint64_t committed_columns, committed_trace_cells;
    const char* native_unit;
    uint64_t native_units;
};

const sj::Value& field(const sj::Value& o, const char* k) {
    const sj::Value* v = o.get(k);
    if (!v) fail("MissingField", std::string("missing field \"") + k + "\"");
    return *v;
}

void exact_keys(const sj::Value& o, std::initializer_list<const char*> names, const char* code) {
    if (o.kind != sj::Value::Object) fail(code, "expected an object");
    if (o.obj.size() != names.size()) fail(code, "unexpected set of fields");
    for (const char* n : names)
        if (!o.get(n)) fail(code, std::string("missing field \"") + n + "\"");
}

const std::string& str_field(const sj::Value& o, const char* k) {
    const sj::Value& v = field(o, k);
    if (v.kind != sj::Value::String) fail("InvalidField", std::string("\"") + k + "\" must be a string");
    return v.s;
}

uint64_t uint_field(const sj::Value& o, const char* k, uint64_t lo, uint64_t hi, const char* code = "InvalidField") {
    const sj::Value& v = field(o, k);
    if (!v.is_uint() || v.mag < lo || v.mag > hi)
        fail(code, std::string("\"") + k + "\" must be an integer in " + std::to_string(lo) + ".." +
                       std::to_string(hi));
    return v.mag;
}

void expect_header(const sj::Value& o, const char* type) {
    if (str_field(o, "protocol") != kProtocol) fail("InvalidField", "unknown protocol");
    const sj::Value& v = field(o, "version");
    if (!v.is_uint() || v.mag != kVersion) fail("InvalidField", "unsupported protocol version");
    if (str_field(o, "type") != type) fail("InvalidField", "unexpected frame type");
}

bool valid_request_id(const std::string& s) {  // protocol.zig:304-309
    if (s.empty() || s.size() > 128 || !isalnum((unsigned char)s[0])) return false;
    for (char c : s)
        if (!isalnum((unsigned char)c) && c != '.' && c != '_' && c != ':' && c != '-') return false;
    return true;
}

std::string parent_dir(const std::string& p) {
    size_t i = p.rfind('/');
    return i == 0 ? "/" : p.substr(0, i);
}

void check_output_path(const std::string& p) {  // protocol.zig:181-196
    if (p.empty() || p[0] != '/') fail("InvalidPath", "output paths must be absolute");
    if (p.back() == '/') fail("InvalidPath", "output path names a directory");
    struct stat st;
    if (lstat(p.c_str(), &st) == 0) fail("OutputAlreadyExists", "output exists: " + p);
    if (errno != ENOENT) fail("InvalidPath", "cannot check output " + p);
    std::string d = parent_dir(p);
    if (stat(d.c_str(), &st) != 0 || !S_ISDIR(st.st_mode)) fail("InvalidPath", "no directory for output " + p);
}

Workload parse_workload(const sj::Value& v) {
    if (v.kind != sj::Value::Object) fail("InvalidWorkload", "workload must be an object");
    const std::string& ex = str_field(v, "example");
    Workload w;
    w.example = ex;
    // Bounds: the shapes the device prover accepts (each prove call checks them again).
    if (ex == "wide_fibonacci") {
        exact_keys(v, {"example", "log_n_rows", "sequence_len"}, "InvalidWorkload");
        w.log_rows = (uint32_t)uint_field(v, "log_n_rows", 2, 26, "InvalidWorkload");
        w.width = (uint32_t)uint_field(v, "sequence_len", 3, 4096, "InvalidWorkload");
    } else if (ex == "poseidon") {
        exact_keys(v, {"example", "log_n_instances"}, "InvalidWorkload");
        w.log_rows = (uint32_t)uint_field(v, "log_n_instances", 4, 29, "InvalidWorkload");
    } else if (ex == "xor") {
        exact_keys(v, {"example", "log_size", "log_step", "offset"}, "InvalidWorkload");
        w.log_rows = (uint32_t)uint_field(v, "log_size", 2, 26, "InvalidWorkload");
        w.width = (uint32_t)uint_field(v, "log_step", 0, w.log_rows, "InvalidWorkload");
        w.value = uint_field(v, "offset", 0, UINT64_MAX, "InvalidWorkload");
    } else if (ex == "xor_logup") {
        exact_keys(v, {"example", "log_n", "seed"}, "InvalidWorkload");
        w.log_rows = (uint32_t)uint_field(v, "log_n", 16, 26, "InvalidWorkload");
        w.value = uint_field(v, "seed", 0, UINT64_MAX, "InvalidWorkload");
    } else {
        fail("InvalidWorkload", "unknown example \"" + ex + "\"");
    }
    return w;
}

ProofProtocol parse_proof_protocol(const std::string& name) {  // native/config.zig:140-151
    ProofProtocol p;
    p.name = name;
    if (name == "smoke") {
        p.pow_bits = 0;
    } else if (name == "functional") {
        p.pow_bits = 10;
    } else if (name == "secure") {
        p.pow_bits = 26;
        p.n_queries = 70;
    } else {
        fail("InvalidProofProtocol", "proof_protocol must be smoke, functional or secure");
    }
    return p;
}

Request parse_request(const sj::Value& o, uint64_t expected_sequence) {  // protocol.zig:77-166
    exact_keys(o,
               {"protocol", "version", "type", "sequence", "request_id", "workload", "proof_protocol", "warmups",
                "samples", "outputs", "budget_gib"},
               "InvalidFields");
    expect_header(o, "prove");
    Request r;
    const sj::Value& seq = field(o, "sequence");
    if (!seq.is_uint() || seq.mag != expected_sequence)
        fail("InvalidSequence", "expected sequence " + std::to_string(expected_sequence));
    r.sequence = seq.mag;
    r.request_id = str_field(o, "request_id");
    if (!valid_request_id(r.request_id)) fail("InvalidRequestId", "request_id: 1-128 of [A-Za-z0-9._:-], alphanumeric first");
    r.budget_gib = str_field(o, "budget_gib");
    {
        char* endp = nullptr;
        errno = 0;
        double g = r.budget_gib.empty() ? 0 : strtod(r.budget_gib.c_str(), &endp);
        if (r.budget_gib.empty() || errno || *endp || !(g > 0) || g > 1e6) fail("InvalidBudget", "budget_gib must be a positive number");
        r.budget_bytes = (uint64_t)(g * 1073741824.0);
        if (!r.budget_bytes) fail("InvalidBudget", "budget_gib below one byte");
    }
    r.w = parse_workload(field(o, "workload"));
    r.proto = parse_proof_protocol(str_field(o, "proof_protocol"));
    r.warmups = (uint32_t)uint_field(o, "warmups", 0, kMaxWarmups);
    r.samples = (uint32_t)uint_field(o, "samples", 1, kMaxSamples);
    const sj::Value& out = field(o, "outputs");
    exact_keys(out, {"proof", "report"}, "InvalidOutputs");
    r.proof_out = str_field(out, "proof");
    r.report_out = str_field(out, "report");
    if (r.proof_out == r.report_out) fail("InvalidOutputs", "proof and report outputs must differ");
    check_output_path(r.proof_out);
    check_output_path(r.report_out);
    return r;
}

void parse_shutdown(const sj::Value& o, uint64_t next_sequence) {  // protocol.zig:168-179
    exact_keys(o, {"protocol", "version", "type", "next_sequence"}, "InvalidShutdown");
    expect_header(o, "shutdown");
    const sj::Value& v = field(o, "next_sequence");
    if (!v.is_uint() || v.mag != next_sequence)
        fail("InvalidSequence", "shutdown must name the next sequence " + std::to_string(next_sequence));
}

Geometry geometry(const Workload& w) {
    Geometry g{};
    g.committed_trees = 2;
    if (w.example == "poseidon") {
        g.trace_log_rows = w.log_rows - 3;  // 8 instances per row (examples/poseidon.zig)
        g.committed_columns = 1264;
        g.native_unit = "poseidon_instances";
        g.native_units = 1ull << w.log_rows;
    } else {
        g.trace_log_rows = w.log_rows;
        if (w.example == "wide_fibonacci") {
            g.committed_columns = w.width;
            g.native_unit = "trace_rows";
        } else if (w.example == "xor") {
            g.committed_columns = 3;  // is_first, is_step (preprocessed) and the main column
            g.native_unit = "xor_rows";
        } else {
            g.committed_trees = 3;    // preprocessed, main, interaction
            g.committed_columns = 12;  // 4 + 4 + the 4 coordinates of the cumulative column
            g.native_unit = "lookups";
        }
        g.native_units = 1ull << w.log_rows;
    }
    g.trace_rows = 1ull << g.trace_log_rows;
    g.committed_trace_cells = g.trace_rows * g.committed_columns;
    return g;
}

std::string descriptor(const Workload& w, const ProofProtocol& p) {  // native/examples.zig:125-175
    char b[512];
    int n;
    if (w.example == "wide_fibonacci")
        n = snprintf(b, sizeof b, "native-proof-workload-v3|example=wide_fibonacci|log_n_rows=%u|sequence_len=%u",
                     w.log_rows, w.width);
    else if (w.example == "poseidon")
        n = snprintf(b, sizeof b, "native-proof-workload-v3|example=poseidon|log_n_instances=%u", w.log_rows);
    else if (w.example == "xor")
        n = snprintf(b, sizeof b, "native-proof-workload-v3|example=xor|log_size=%u|log_step=%u|offset=%llu",
                     w.log_rows, w.width, (unsigned long long)w.value);
    else
        n = snprintf(b, sizeof b, "native-proof-workload-v3|example=xor_logup|log_n=%u|seed=%llu", w.log_rows,
                     (unsigned long long)w.value);
    snprintf(b + n, sizeof b - n,
             "|protocol=%s|pow_bits=%u|log_blowup_factor=%u|log_last_layer_degree_bound=%u|n_queries=%u|fold_step=%u",
             p.name.c_str(), p.pow_bits, p.log_blowup_factor, p.log_last_layer_degree_bound, p.n_queries, p.fold_step);
    return b;
}

stwo_rocm_serve_job job_of(const Workload& w) {
    stwo_rocm_serve_job j{};
    if (w.example == "wide_fibonacci") j.kind = 0, j.seq_len = w.width;
    else if (w.example == "poseidon") j.kind = 2;
    else if (w.example == "xor") j.kind = 3, j.seq_len = w.width;
    else j.kind = 4;
    j.log_n = w.log_rows;
    return j;
}

std::string workload_parameters(const Workload& w) {
    sj::Obj o;
    if (w.example == "wide_fibonacci") o.u("log_n_rows", w.log_rows).u("sequence_len", w.width);
    else if (w.example == "poseidon") o.u("log_n_instances", w.log_rows);
    else if (w.example == "xor") o.u("log_size", w.log_rows).u("log_step", w.width).u("offset", w.value);
    else o.u("log_n", w.log_rows).u("seed", w.value);
    return o.done();
}

// The statement a verifier needs: the public part of the workload (the XOR-lookup seed only
// chooses the witness) and the PCS configuration.
std::string statement_json(const Request& r) {
    const Workload& w = r.w;
    std::string st = w.example == "xor_logup" ? sj::Obj().u("log_n", w.log_rows).done() : workload_parameters(w);
    return sj::Obj()
        .str("air", w.example)
        .raw("statement", st)
        .raw("pcs_config", sj::Obj()
                               .u("pow_bits", r.proto.pow_bits)
                               .u("log_blowup_factor", r.proto.log_blowup_factor)
                               .u("log_last_layer_degree_bound", r.proto.log_last_layer_degree_bound)
                               .u("n_queries", r.proto.n_queries)
                               .done())
        .done();
}

// ------------------------------------------------------------------ statistics.zig:10-36
struct Summary {
    double median, min, max, mad;
};
double middle(const std::vector<double>& v) {
    size_t u = v.size() / 2;
    return (v.size() & 1) ? v[u] : (v[u - 1] + v[u]) / 2.0;
}
Summary summarize(std::vector<double> v) {
    std::sort(v.begin(), v.end());
    Summary s;
    s.median = middle(v);
    s.min = v.front();
    s.max = v.back();
    std::vector<double> dev;
    for (double x : v) dev.push_back(std::fabs(x - s.median));
    std::sort(dev.begin(), dev.end());
    s.mad = middle(dev);
    return s;
}
std::string summary_json(const Summary& s) {
    return sj::Obj().f("median", s.median).f("min", s.min).f("max", s.max).f("mad", s.mad).done();
}

// ------------------------------------------------------------------ files
std::string file_sha256(const std::string& path) {
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) return "";
    sj::Sha256 h;
    std::vector<char> b(1 << 16);
    size_t n;
    while ((n = fread(b.data(), 1, b.size(), f)) > 0) h.update(b.data(), n);
    fclose(f);
    uint8_t d[32];
    h.finish(d);
    return sj::hex(d, 32);
}

void write_new_file(const std::string& path, const void* data, size_t n) {
    int fd = open(path.c_str(), O_WRONLY | O_CREAT | O_EXCL | O_CLOEXEC, 0644);
    if (fd < 0) fail("OutputWriteFailed", "cannot create " + path);
    const char* p = (const char*)data;
    while (n) {
        ssize_t k = write(fd, p, n);
        if (k < 0 && errno == EINTR) continue;
        if (k <= 0) {
            close(fd);
            unlink(path.c_str());
            fail("OutputWriteFailed", "cannot write " + path);
        }
        p += k;
        n -= (size_t)k;
    }
    if (close(fd) != 0) fail("OutputWriteFailed", "cannot write " + path);
}

// Files written aside for one request; whatever is not published is removed.
struct TempFiles {
    std::vector<std::string> paths;
    ~TempFiles() {
        for (auto& p : paths) unlink(p.c_str());
    }
    std::string make(const std::string& final_path, uint64_t seq, const char* tag) {
        std::string p = final_path + ".tmp-" + std::to_string((long)getpid()) + "-" + std::to_string(seq) + "-" + tag;
        paths.push_back(p);
        return p;
    }
};

// ------------------------------------------------------------------ verifier
struct Verifier {
    std::string path;
    std::vector<std::string> args;
    std::string sha256;
    bool configured() const { return !path.empty(); }
};

// Runs the verifier as a child process, its stdout sent to stderr (stdout carries the frames).
bool run_verifier(const Verifier& v, const std::string& statement, const std::string& proof) {
    std::vector<std::string> a = {v.path};
    a.insert(a.end(), v.args.begin(), v.args.end());
    a.insert(a.end(), {"--statement", statement, "--proof", proof});
    std::vector<char*> argv;
    for (auto& s : a) argv.push_back(const_cast<char*>(s.c_str()));
    argv.push_back(nullptr);
    posix_spawn_file_actions_t fa;
    posix_spawn_file_actions_init(&fa);
    posix_spawn_file_actions_addopen(&fa, 0, "/dev/null", O_RDONLY, 0);
    posix_spawn_file_actions_adddup2(&fa, 2, 1);
    pid_t pid;
    int rc = posix_spawn(&pid, v.path.c_str(), &fa, nullptr, argv.data(), environ);
    posix_spawn_file_actions_destroy(&fa);
    if (rc != 0) fail("VerifierFailed", "cannot start the verifier " + v.path);
    int status = 0;
    while (waitpid(pid, &status, 0) < 0) {
        if (errno != EINTR) fail("VerifierFailed", "lost the verifier process");
    }
    return WIFEXITED(status) && WEXITSTATUS(status) == 0;
}

// ------------------------------------------------------------------ output frames
void emit(const std::string& frame) {
    fwrite(frame.data(), 1, frame.size(), stdout);
    fputc('\n', stdout);
    fflush(stdout);
}

sj::Obj header(const char* type) {
    sj::Obj o;
    o.str("protocol", kProtocol).u("version", kVersion).str("type", type);
    return o;
}

// Reads one frame (a line without its newline); false at the end of input.
bool read_frame(std::string& line) {
    line.clear();
    int c;
    bool any = false;
    while ((c = getchar_unlocked()) != EOF) {
        any = true;
        if (c == '\n') return true;
        if (line.size() >= kMaxFrame) fail("FrameTooLarge", "frames are limited to 4 MiB");
        line += (char)c;
    }
    return any;
}

std::vector<std::pair<std::string, std::string>> environment_overrides() {
    // Variables that change how kernels run or how this library behaves.
    static const char* watched[] = {"HIP_LAUNCH_BLOCKING", "AMD_SERIALIZE_KERNEL", "AMD_SERIALIZE_COPY",
                                    "AMD_LOG_LEVEL", "HSA_XNACK", "HIP_FORCE_DEV_KERNARG"};
    std::vector<std::pair<std::string, std::string>> out;
    for (char** e = environ; *e; e++) {
        std::string kv = *e;
        size_t eq = kv.find('=');
        if (eq == std::string::npos) continue;
        std::string k = kv.substr(0, eq);
        bool hit = k.rfind("STWO_ROCM_", 0) == 0;
        for (const char* w : watched) hit = hit || k == w;
        if (hit) out.emplace_back(k, kv.substr(eq + 1));
    }
    std::sort(out.begin(), out.end());
    return out;
}

uint64_t host_peak_rss() {
    struct rusage ru;
    return getrusage(RUSAGE_SELF, &ru) == 0 ? (uint64_t)ru.ru_maxrss * 1024 : 0;
}

std::string library_path() {
    Dl_info info;
    if (dladdr((void*)&stwo_rocm_ctx_create, &info) && info.dli_fname) return info.dli_fname;
    return "";
}

// ------------------------------------------------------------------ the session
struct Session {
    bool check_only = false;
    int device = 0;
    Verifier verifier;
    stwo_rocm_ctx* ctx = nullptr;
    std::string identity;  // name;arch;CUs;bytes
    std::string session_id;
    double init_seconds = 0;
    std::string exe_sha256, lib_path, lib_sha256;
    std::vector<uint8_t> buf = std::vector<uint8_t>(4u << 20);

    ~Session() {
        if (ctx) stwo_rocm_ctx_destroy(ctx);
    }

    void lib_check(int rc, const char* what) {
        if (rc == STWO_ROCM_OK) return;
        std::string m = std::string(what) + ": " + (ctx ? stwo_rocm_last_error(ctx) : "error");
        fail(rc == STWO_ROCM_EBUDGET ? "DeviceBudgetExceeded" : "ProveFailed", m);
    }

    std::vector<std::string> identity_parts() const {
        std::vector<std::string> p;
        size_t s = 0;
        for (;;) {
            size_t e = identity.find(';', s);
            p.push_back(identity.substr(s, e == std::string::npos ? std::string::npos : e - s));
            if (e == std::string::npos) break;
            s = e + 1;
        }
        while (p.size() < 4) p.push_back("");
        return p;
    }

    void start() {
        session_id = "rocm-" + std::to_string((long)getpid()) + "-" +
                     std::to_string((long long)std::chrono::system_clock::now().time_since_epoch().count());
        exe_sha256 = file_sha256("/proc/self/exe");
        lib_path = library_path();
        lib_sha256 = lib_path.empty() ? "" : file_sha256(lib_path);
        if (verifier.configured()) {
            if (verifier.path[0] != '/' || access(verifier.path.c_str(), X_OK) != 0)
                fail("VerifierUnavailable", "the verifier must be an absolute path to an executable");
            verifier.sha256 = file_sha256(verifier.path);
        }
        sj::Obj dev;
        if (!check_only) {
            double t0 = now_s();
            char err[512] = {0};
            if (stwo_rocm_ctx_create(device, 0, &ctx, err, sizeof err) != STWO_ROCM_OK)
                fail("DeviceUnavailable", std::string("device context: ") + err);
            char id[512] = {0};
            lib_check(stwo_rocm_ctx_identity(ctx, id, sizeof id), "device identity");
            identity = id;
            init_seconds = now_s() - t0;
            auto p = identity_parts();
            dev.u("ordinal", (uint64_t)device)
                .str("name", p[0])
                .str("arch", p[1])
                .u("compute_units", strtoull(p[2].c_str(), nullptr, 10))
                .u("memory_bytes", strtoull(p[3].c_str(), nullptr, 10));
        }
        emit(header("ready")
                 .str("session_id", session_id)
                 .str("mode", check_only ? "check" : "prove")
                 .raw("device", check_only ? "null" : dev.done())
                 .str("session_executable_sha256", exe_sha256)
                 .str("library_sha256", lib_sha256)
                 .raw("verifier", sj::Obj()
                                      .b("required", verifier.configured())
                                      .raw("executable", verifier.configured() ? sj::quote(verifier.path) : "null")
                                      .raw("executable_sha256",
                                           verifier.configured() ? sj::quote(verifier.sha256) : "null")
                                      .done())
                 .raw("capabilities", sj::Obj()
                                          .b("strict_order", true)
                                          .b("atomic_outputs", true)
                                          .b("verified_proofs", verifier.configured())
                                          .b("runtime_reuse", true)
                                          .b("resident_arena_reuse", true)
                                          .b("twiddle_tower_reuse", true)
                                          .done())
                 .done());
    }

    // One proof through the C-ABI; grows the output buffer when the proof does not fit (the
    // call that fits is the one timed).
    double prove_once(const Request& r, uint64_t& len, stwo_rocm_prove_stats* st) {
        const Workload& w = r.w;
        const ProofProtocol& p = r.proto;
        for (int attempt = 0; attempt < 2; attempt++) {
            len = 0;
            double t0 = now_s();
            int rc;
            if (w.example == "wide_fibonacci") {
                stwo_rocm_widefib_params q{w.log_rows, w.width, p.pow_bits, p.log_blowup_factor,
                                           p.log_last_layer_degree_bound, p.n_queries, 0, 0};
                rc = stwo_rocm_prove_wide_fibonacci(ctx, &q, nullptr, buf.data(), buf.size(), &len, st);
            } else if (w.example == "poseidon") {
                stwo_rocm_poseidon_params q{w.log_rows, p.pow_bits, p.log_blowup_factor, p.log_last_layer_degree_bound,
                                            p.n_queries};
                rc = stwo_rocm_prove_poseidon(ctx, &q, buf.data(), buf.size(), &len, st);
            } else if (w.example == "xor") {
                stwo_rocm_xor_params q{w.log_rows, w.width, w.value, p.pow_bits, p.log_blowup_factor,
                                       p.log_last_layer_degree_bound, p.n_queries};
                rc = stwo_rocm_prove_xor(ctx, &q, buf.data(), buf.size(), &len, st);
            } else {
                stwo_rocm_logup_params q{w.log_rows, p.pow_bits, w.value, p.log_blowup_factor,
                                         p.log_last_layer_degree_bound, p.n_queries, 0};
                rc = stwo_rocm_prove_logup_xor(ctx, &q, buf.data(), buf.size(), &len);
            }
            double dt = now_s() - t0;
            if (rc == STWO_ROCM_EINVAL && len > buf.size() && attempt == 0) {
                buf.resize(len + (len >> 3));
                continue;
            }
            lib_check(rc, "prove");
            return dt;
        }
        fail("ProveFailed", "proof buffer");
    }

    struct SampleRecord {
        double input_s = 0, prove_s = 0, encode_s = 0, verify_s = 0, request_s = 0;
        uint64_t bytes = 0;
        std::string sha256;
        bool verified = false;
        bool has_stats = false;
        stwo_rocm_prove_stats st{};
    };

    // One verified proof: prove, hash, and (with a verifier) write it aside and check it.
    SampleRecord one_proof(const Request& r, const std::string& statement_path,
                           const std::string& proof_path, std::vector<uint8_t>* keep) {
        SampleRecord s;
        bool stats = r.w.example != "xor_logup";
        double t0 = now_s();
        uint64_t len = 0;
        double call_s = prove_once(r, len, stats ? &s.st : nullptr);
        s.has_stats = stats;
        s.encode_s = stats ? s.st.stage_host_ms[7] / 1e3 : 0.0;
        s.prove_s = call_s - s.encode_s;
        s.bytes = len;
        s.sha256 = sj::sha256_hex(buf.data(), len);
        if (keep) keep->assign(buf.begin(), buf.begin() + len);
        if (verifier.configured()) {
            double tv = now_s();
            unlink(proof_path.c_str());
            write_new_file(proof_path, buf.data(), len);
            s.verified = run_verifier(verifier, statement_path, proof_path);
            s.verify_s = now_s() - tv;
            if (!s.verified) fail("ProofRejected", "the verifier rejected the proof of request " + r.request_id);
        }
        s.request_s = now_s() - t0;
        return s;
    }

    void serve_request(const Request& r) {
        const Geometry g = geometry(r.w);
        const std::string desc = descriptor(r.w, r.proto);
        const std::string desc_sha = sj::sha256_hex(desc.data(), desc.size());
        stwo_rocm_serve_job job = job_of(r.w);
        uint64_t accounted = 0;
        if (stwo_rocm_job_device_bytes(&job, &accounted) != STWO_ROCM_OK)
            fail("InvalidWorkload", "no device-byte estimate for this workload");
        if (accounted > r.budget_bytes)
            fail("AccountedMemoryBudgetExceeded", "the workload needs an estimated " + std::to_string(accounted) +
                                                      " device bytes, over the budget of " +
                                                      std::to_string(r.budget_bytes));
        if (check_only) {
            emit(header("checked")
                     .u("sequence", r.sequence)
                     .str("request_id", r.request_id)
                     .str("descriptor", desc)
                     .str("descriptor_sha256", desc_sha)
                     .u("accounted_bytes", accounted)
                     .u("budget_bytes", r.budget_bytes)
                     .b("admitted", true)
                     .done());
            return;
        }
        const double t_block = now_s();
        stwo_rocm_runtime_info before{}, after{};
        lib_check(stwo_rocm_get_runtime_info(ctx, 1, &before), "runtime info");
        // The budget holds for this request: cached blocks beyond it are released now, and an
        // allocation that would take the live bytes over it fails the request.
        lib_check(stwo_rocm_set_budget(ctx, r.budget_bytes), "budget");

        TempFiles tmp;
        std::string statement_path, proof_check_path;
        if (verifier.configured()) {
            statement_path = tmp.make(r.proof_out, r.sequence, "statement.json");
            proof_check_path = tmp.make(r.proof_out, r.sequence, "check");
            std::string st = statement_json(r);
            write_new_file(statement_path, st.data(), st.size());
        }
        std::vector<double> warm_s;
        size_t verified = 0;
        for (uint32_t i = 0; i < r.warmups; i++) {
            SampleRecord s = one_proof(r, statement_path, proof_check_path, nullptr);
            warm_s.push_back(s.request_s);
            verified += s.verified;
        }
        std::vector<SampleRecord> samples;
        std::vector<uint8_t> proof;
        bool identical = true;
        for (uint32_t i = 0; i < r.samples; i++) {
            SampleRecord s = one_proof(r, statement_path, proof_check_path, i == 0 ? &proof : nullptr);
            if (i > 0) identical = identical && s.sha256 == samples[0].sha256 && s.bytes == samples[0].bytes &&
                                   memcmp(buf.data(), proof.data(), proof.size()) == 0;
            verified += s.verified;
            samples.push_back(s);
        }
        if (!identical) fail("ProofNotDeterministic", "samples of one request produced different proofs");
        lib_check(stwo_rocm_get_runtime_info(ctx, 0, &after), "runtime info");
        const size_t verified_samples = verifier.configured() ? r.samples : 0;

        // ---- report (native/report.zig:286-304 layout)
        std::vector<double> in_v, pr_v, en_v, ve_v, rq_v, nat, rnat, row, rrow, cell;
        sj::Arr sample_arr, proof_arr, tele_arr;
        for (auto& s : samples) {
            in_v.push_back(s.input_s);
            pr_v.push_back(s.prove_s);
            en_v.push_back(s.encode_s);
            ve_v.push_back(s.verify_s);
            rq_v.push_back(s.request_s);
            nat.push_back(g.native_units / s.prove_s / 1e6);
            rnat.push_back(g.native_units / s.request_s / 1e6);
            row.push_back(g.trace_rows / s.prove_s / 1e6);
            rrow.push_back(g.trace_rows / s.request_s / 1e6);
            cell.push_back(g.committed_trace_cells / s.prove_s / 1e6);
            sample_arr.raw(sj::Obj()
                               .f("input_seconds", s.input_s)
                               .f("prove_seconds", s.prove_s)
                               .f("proof_encode_seconds", s.encode_s)
                               .f("verify_seconds", s.verify_s)
                               .f("request_seconds", s.request_s)
                               .f("native_mhz", nat.back())
                               .f("request_native_mhz", rnat.back())
                               .f("trace_row_mhz", row.back())
                               .f("request_trace_row_mhz", rrow.back())
                               .f("committed_mcells_per_second", cell.back())
                               .done());
            proof_arr.raw(sj::Obj().u("bytes", s.bytes).str("sha256", s.sha256).done());
            if (s.has_stats) {
                sj::Arr gpu, host;
                for (int k = 0; k < 8; k++) {
                    gpu.raw(sj::num(s.st.stage_gpu_ms[k]));
                    host.raw(sj::num(s.st.stage_host_ms[k]));
                }
                tele_arr.raw(sj::Obj()
                                 .raw("stage_gpu_ms", gpu.done())
                                 .raw("stage_host_ms", host.done())
                                 .f("total_gpu_ms", s.st.total_gpu_ms)
                                 .f("total_host_ms", s.st.total_host_ms)
                                 .u("peak_device_bytes", s.st.peak_hbm_bytes)
                                 .done());
            }
        }
        const Summary prove_sum = summarize(pr_v);
        const size_t min_samples = prove_sum.median < 1.0 ? 5 : 3;  // runner.zig:392
        const bool sampling = r.warmups >= kMinHeadlineWarmups && r.samples >= min_samples;
        const bool is_verified = verifier.configured();
        const char* evidence = !is_verified ? "unverified" : sampling ? "verified_unprofiled" : "correctness_only";
        const auto overrides = environment_overrides();
        bool telemetry_valid = true;  // every sample carries its stage timings (none are kept for XOR-lookup)
        for (auto& s : samples) telemetry_valid = telemetry_valid && (!s.has_stats || s.st.total_gpu_ms > 0);
#ifdef NDEBUG
        const bool release = true;
#else
        const bool release = false;
#endif
        const bool headline_ok = is_verified && sampling && r.proto.name != "smoke" && release && overrides.empty() &&
                                 identical && telemetry_valid;
        sj::Arr ov;
        for (auto& kv : overrides) ov.raw(sj::Obj().str("name", kv.first).str("value", kv.second).done());
        auto parts = identity_parts();
        struct utsname un;
        uname(&un);
        auto headline = [&](const std::vector<double>& v) {
            return headline_ok ? summary_json(summarize(v)) : std::string("null");
        };
        sj::Arr warm_arr;
        for (double x : warm_s) warm_arr.raw(sj::num(x));
        const uint64_t request_peak = std::max(after.peak_live, before.bytes_live);
        std::string report =
            sj::Obj()
                .u("schema_version", 7)
                .str("backend", "rocm_hip")
                .str("evidence_class", evidence)
                .b("profiled", false)
                .raw("provenance", sj::Obj()
                                       .str("git_commit", STWO_ROCM_GIT_COMMIT)
                                       .b("git_dirty", STWO_ROCM_GIT_DIRTY != 0)
                                       .str("session_executable_sha256", exe_sha256)
                                       .str("library_path", lib_path)
                                       .str("library_sha256", lib_sha256)
                                       .str("compiler", __VERSION__)
                                       .str("optimization", release ? "release" : "debug")
                                       .str("target_os", un.sysname)
                                       .str("target_arch", un.machine)
                                       .str("device_name", parts[0])
                                       .str("device_arch", parts[1])
                                       .u("compute_units", strtoull(parts[2].c_str(), nullptr, 10))
                                       .u("device_memory_bytes", strtoull(parts[3].c_str(), nullptr, 10))
                                       .u("cpu_count", (uint64_t)sysconf(_SC_NPROCESSORS_ONLN))
                                       .raw("environment_overrides", ov.done())
                                       .b("complete", overrides.empty() && !lib_sha256.empty() &&
                                                          strcmp(STWO_ROCM_GIT_COMMIT, "unknown") != 0)
                                       .done())
                .raw("protocol", sj::Obj()
                                     .str("name", r.proto.name)
                                     .u("pow_bits", r.proto.pow_bits)
                                     .u("log_blowup_factor", r.proto.log_blowup_factor)
                                     .u("log_last_layer_degree_bound", r.proto.log_last_layer_degree_bound)
                                     .u("n_queries", r.proto.n_queries)
                                     .u("fold_step", r.proto.fold_step)
                                     .done())
                .raw("workload", sj::Obj()
                                     .str("name", r.w.example)
                                     .str("descriptor_sha256", desc_sha)
                                     .raw("parameters", workload_parameters(r.w))
                                     .u("trace_log_rows", g.trace_log_rows)
                                     .u("trace_rows", g.trace_rows)
                                     .u("committed_trees", g.committed_trees)
                                     .u("committed_columns", g.committed_columns)
                                     .u("committed_trace_cells", g.committed_trace_cells)
                                     .str("native_unit", g.native_unit)
                                     .u("native_units", g.native_units)
                                     .done())
                .raw("resource_admission", sj::Obj()
                                               .str("profile", "device_budget")
                                               .u("committed_cells", g.committed_trace_cells)
                                               .u("accounted_bytes", accounted)
                                               .u("max_accounted_bytes", r.budget_bytes)
                                               .str("budget_gib", r.budget_gib)
                                               .done())
                .raw("resources", sj::Obj()
                                      .str("measurement_scope", "session_request")
                                      .str("source", "device_arena")
                                      .u("measured_warmups", r.warmups)
                                      .u("measured_samples", r.samples)
                                      .u("request_peak_device_bytes", request_peak)
                                      .u("device_bytes_live_after", after.bytes_live)
                                      .u("device_bytes_cached_after", after.bytes_cached)
                                      .u("lifetime_peak_physical_footprint_bytes", host_peak_rss())
                                      .null("energy_nj")
                                      .null("instructions")
                                      .null("cycles")
                                      .u("canonical_proof_bytes", proof.size())
                                      .b("within_budget", request_peak <= r.budget_bytes)
                                      .b("complete", true)
                                      .null("unavailable_reason")
                                      .done())
                .raw("session", sj::Obj()
                                    .str("session_id", session_id)
                                    .u("max_circle_log", after.tower_log)
                                    .u("device_byte_budget", r.budget_bytes)
                                    .u("retained_device_twiddle_bytes", after.tower_bytes)
                                    .u("tower_build_count", after.tower_builds)
                                    .u("requests_completed_before", r.sequence)
                                    .done())
                .raw("runtime_admission", sj::Obj()
                                              .b("initialized", true)
                                              .str("origin", "hip")
                                              .str("source_sha256", lib_sha256)
                                              .u("active_call_leases", 0)
                                              .u("live_device_bytes", after.bytes_live)
                                              .str("platform_identity", identity)
                                              .u("initialization_count", 1)
                                              .u("shutdown_count", 0)
                                              .done())
                .raw("proof", sj::Obj()
                                  .raw("samples", proof_arr.done())
                                  .u("verified_samples", verified_samples)
                                  .b("all_samples_byte_identical", identical)
                                  .raw("artifact", sj::Obj()
                                                       .str("path", r.proof_out)
                                                       .u("sample_index", 0)
                                                       .u("bytes", proof.size())
                                                       .str("sha256", samples[0].sha256)
                                                       .str("exchange_mode", "canonical_json_wire")
                                                       .done())
                                  .done())
                .raw("backend_telemetry",
                     r.w.example == "xor_logup"
                         ? std::string("null")
                         : sj::Obj()
                               .str("scope", "proof_request")
                               .raw("stage_names", "[\"main_commit\",\"composition\",\"sampled_values\",\"quotients\","
                                                   "\"fri_commit\",\"pow\",\"decommit\",\"encode\"]")
                               .u("warmups", r.warmups)
                               .raw("samples", tele_arr.done())
                               .u("total_cpu_fallbacks", 0)
                               .b("valid", telemetry_valid)
                               .done())
                .raw("timing", sj::Obj()
                                   .f("backend_init_seconds", init_seconds)
                                   .raw("warmup_request_seconds", warm_arr.done())
                                   .raw("samples", sample_arr.done())
                                   .null("stage_profiles")
                                   .str("scope", kTimingScope)
                                   .raw("input_seconds", summary_json(summarize(in_v)))
                                   .raw("prove_seconds", summary_json(prove_sum))
                                   .raw("proof_encode_seconds", summary_json(summarize(en_v)))
                                   .raw("verify_seconds", summary_json(summarize(ve_v)))
                                   .raw("request_seconds", summary_json(summarize(rq_v)))
                                   .done())
                .raw("throughput", sj::Obj()
                                       .b("headline_eligible", headline_ok)
                                       .raw("headline_native_mhz", headline(nat))
                                       .null("diagnostic_native_mhz")
                                       .raw("headline_request_native_mhz", headline(rnat))
                                       .null("diagnostic_request_native_mhz")
                                       .raw("headline_trace_row_mhz", headline(row))
                                       .null("diagnostic_trace_row_mhz")
                                       .raw("headline_request_trace_row_mhz", headline(rrow))
                                       .null("diagnostic_request_trace_row_mhz")
                                       .raw("headline_committed_mcells_per_second", headline(cell))
                                       .null("diagnostic_committed_mcells_per_second")
                                       .raw("headline_requirements",
                                            sj::Obj()
                                                .b("verified_unprofiled", is_verified && sampling)
                                                .b("sampling_contract", sampling)
                                                .b("functional_protocol", r.proto.name != "smoke")
                                                .b("release_build", release)
                                                .b("clean_complete_provenance", overrides.empty())
                                                .b("byte_identical_verified_samples", identical)
                                                .b("backend_telemetry_valid", telemetry_valid)
                                                .done())
                                       .done())
                .done();

        // ---- publish: both files written aside, then linked into place (never overwriting)
        std::string proof_tmp = tmp.make(r.proof_out, r.sequence, "proof");
        std::string report_tmp = tmp.make(r.report_out, r.sequence, "report");
        write_new_file(proof_tmp, proof.data(), proof.size());
        report += "\n";
        write_new_file(report_tmp, report.data(), report.size());
        if (link(proof_tmp.c_str(), r.proof_out.c_str()) != 0)
            fail(errno == EEXIST ? "OutputAlreadyExists" : "OutputWriteFailed", "cannot publish " + r.proof_out);
        if (link(report_tmp.c_str(), r.report_out.c_str()) != 0) {
            int e = errno;
            unlink(r.proof_out.c_str());
            fail(e == EEXIST ? "OutputAlreadyExists" : "OutputWriteFailed", "cannot publish " + r.report_out);
        }
        Summary req_sum = summarize(rq_v);
        emit(header("result")
                 .str("status", is_verified ? "verified" : "unverified")
                 .u("sequence", r.sequence)
                 .str("request_id", r.request_id)
                 .b("proof_verified", is_verified)
                 .u("verified_proofs", verified)
                 .b("outputs_committed", true)
                 .str("evidence_class", evidence)
                 .str("descriptor_sha256", desc_sha)
                 .u("proof_bytes", proof.size())
                 .str("proof_sha256", samples[0].sha256)
                 .f("prove_wall_s", prove_sum.median)
                 .str("prove_timing_scope", kTimingScope)
                 .f("request_wall_s", req_sum.median)
                 .str("native_unit", g.native_unit)
                 .f("native_mhz", g.native_units / prove_sum.median / 1e6)
                 .f("session_block_wall_s", now_s() - t_block)
                 .u("request_peak_device_bytes", request_peak)
                 .raw("reuse", sj::Obj()
                                   .b("runtime", r.sequence > 0)
                                   .b("resident_arena", before.bytes_cached > 0)
                                   .b("twiddle_tower", after.tower_builds == before.tower_builds)
                                   .done())
                 .done());
    }

    int run() {
        uint64_t next = 0;
        std::string line;
        for (;;) {
            if (!read_frame(line)) fail("UncleanEndOfStream", "input ended without a shutdown frame");
            if (line.empty()) fail("EmptyFrame", "empty frame");
            sj::Value v;
            try {
                v = sj::parse(line);
            } catch (const sj::ParseError& e) {
                fail("InvalidFrame", std::string("malformed JSON: ") + e.what());
            }
            const sj::Value* t = v.kind == sj::Value::Object ? v.get("type") : nullptr;
            if (!t || t->kind != sj::Value::String || (t->s != "prove" && t->s != "shutdown"))
                fail("InvalidFrame", "a frame is an object of type prove or shutdown");
            if (t->s == "shutdown") {
                parse_shutdown(v, next);
                emit(header("closed").u("completed", next).done());
                return 0;
            }
            Request r = parse_request(v, next);
            cur_sequence = (int64_t)r.sequence;
            cur_request = r.request_id;
            serve_request(r);
            cur_sequence = -1;
            cur_request.clear();
            next++;
        }
    }

    int64_t cur_sequence = -1;
    std::string cur_request;
};

void usage() {
    fprintf(stderr,
            "usage: stwo_rocm_session [--device N] [--check] [--verifier ABS_PATH [--verifier-arg ARG]...]\n"
            "JSONL frames on stdin (prove, shutdown), answers on stdout (ready, result, closed, error).\n");
}

}  // namespace

int main(int argc, char** argv) {
    Session s;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        if (a == "--check") {
            s.check_only = true;
        } else if (a == "--device" && i + 1 < argc) {
            s.device = atoi(argv[++i]);
        } else if (a == "--verifier" && i + 1 < argc) {
            s.verifier.path = argv[++i];
        } else if (a == "--verifier-arg" && i + 1 < argc) {
            s.verifier.args.push_back(argv[++i]);
        } else if (a == "--help" || a == "-h") {
            usage();
            return 0;
        } else {
            usage();
            return 2;
        }
    }
    if (!s.verifier.args.empty() && !s.verifier.configured()) {
        usage();
        return 2;
    }
    try {
        s.start();
        return s.run();
    } catch (const Fail& f) {
        sj::Obj o = header("error");
        if (s.cur_sequence >= 0) o.u("sequence", (uint64_t)s.cur_sequence).str("request_id", s.cur_request);
        else o.null("sequence").null("request_id");
        emit(o.str("code", f.code).str("message", f.msg).done());
        return 1;
    } catch (const std::exception& e) {
        emit(header("error").null("sequence").null("request_id").str("code", "Internal").str("message", e.what()).done());
        return 1;
    }
}
