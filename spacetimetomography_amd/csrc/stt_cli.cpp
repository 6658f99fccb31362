// stt_cli.cpp — headless command-line driver of the space-time reconstruction.
//
// Replaces the reference's main.cpp and its GLUT window
// (GLUTSceneContainer.cpp:70-96, whose display callback runs
// ST_Tomography::Render once).  Only the C-ABI of include/stt.h is used.
//
// Flags: the argtable2 table of main.cpp:13-47 with the code defaults of
// main.cpp:67-95, parsed with argtable2's rules (getopt_long underneath):
// "-s 0.3", "-s0.3", "--Sigma=0.3", "--Sigma 0.3", unambiguous long-option
// prefixes, repeated multi-value options ("-f 128 -f 128 -f 64"), integers
// in decimal/0x/0o/0b with KB/MB/GB suffixes; options with a minimum count
// of one (-f, -u, -i, --oxyz, --sdg) are required, as there.
//
// Run: ST_Tomography::Render (ST_Tomography.cpp:397-614).  Each outer
// iteration computes the flow of every frame pair (L1Huber_OF_MultiScale3D,
// :421-469) and saves the second frame warped back by it (TIFF) and the
// flow (Analyze); from the second outer iteration on, primal-dual passes
// over all frames follow (:475-608), and every frame's volume is saved after
// its SART step (:786-805).
//
// Profiles:
//   --compat=reference  the values the reference actually runs with: PD
//       steps sigma = 0.1125, tau = 1 whatever -s/-t say (ST_Tomography.cpp:
//       99-102), the ATV prior whatever --prior says (:93), 20 outer
//       iterations (:95) with 2 PD passes (:477) whatever -b says, one SART
//       sweep whatever -p says (:633), start angles from lds.txt whatever
//       --sdg says (:258-267) and the rose ROI cut[] (:114).
//   default             the flags as main.cpp documents them: -s/-t are the
//       PD steps, --prior the prior, -b the PD passes per outer iteration,
//       -p the SART sweeps per PD step, --sdg the start angles (nframes x
//       nrounds values; else lds.txt); ROI = the whole interior.
// The reference shuffles each SART sweep's views with a time seed (:635-642);
// both profiles sweep the views in table order.
#include <strings.h>
#include <sys/stat.h>
#include <sys/types.h>

#include <cerrno>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "../../include/stt.h"

namespace {

// ------------------------------------------------------------ arguments --

enum class Kind { Dbl, Int, Str, File, Lit };

struct Opt {
    char sh;           // short name or 0
    const char* lng;   // long name or nullptr
    Kind kind;
    int minc, maxc;
    const char* datatype;
    const char* gloss;
    std::vector<std::string> vals;
    int count = 0;
};

struct Args {
    std::deque<Opt> opts;  // stable addresses: main keeps references to entries
    std::vector<std::string> errors;

    Opt& add(char sh, const char* lng, Kind k, int minc, int maxc, const char* dt, const char* gloss) {
        opts.push_back(Opt{sh, lng, k, minc, maxc, dt, gloss, {}, 0});
        return opts.back();
    }
    Opt* by_short(char c) {
        for (Opt& o : opts)
            if (o.sh == c) return &o;
        return nullptr;
    }
    // getopt_long: the first exact match wins, else a unique prefix.
    Opt* by_long(const std::string& name, bool& ambiguous) {
        ambiguous = false;
        for (Opt& o : opts)
            if (o.lng && name == o.lng) return &o;
        Opt* hit = nullptr;
        for (Opt& o : opts)
            if (o.lng && std::strncmp(o.lng, name.c_str(), name.size()) == 0) {
                if (hit && std::strcmp(hit->lng, o.lng) != 0) {
                    ambiguous = true;
                    return nullptr;
                }
                hit = &o;
            }
        return hit;
    }
};

// argtable2's integer syntax: optional sign, decimal / 0x / 0o / 0b digits,
// optional KB / MB / GB suffix.
bool parse_int(const std::string& s, long& out) {
    const char* p = s.c_str();
    bool neg = false;
    if (*p == '+' || *p == '-') neg = *p++ == '-';
    int base = 10;
    if (p[0] == '0' && (p[1] == 'x' || p[1] == 'X')) base = 16, p += 2;
    else if (p[0] == '0' && (p[1] == 'o' || p[1] == 'O')) base = 8, p += 2;
    else if (p[0] == '0' && (p[1] == 'b' || p[1] == 'B')) base = 2, p += 2;
    if (!*p) return false;
    char* end = nullptr;
    errno = 0;
    long v = std::strtol(p, &end, base);
    if (errno || end == p) return false;
    if (!strcasecmp(end, "KB")) v *= 1024L, end += 2;
    else if (!strcasecmp(end, "MB")) v *= 1024L * 1024L, end += 2;
    else if (!strcasecmp(end, "GB")) v *= 1024L * 1024L * 1024L, end += 2;
    if (*end) return false;
    out = neg ? -v : v;
    return true;
}

bool parse_dbl(const std::string& s, double& out) {
    if (s.empty()) return false;
    char* end = nullptr;
    errno = 0;
    out = std::strtod(s.c_str(), &end);
    return !errno && end && *end == 0;
}

std::string opt_name(const Opt& o) {
    if (o.sh && o.lng) return std::string("-") + o.sh + "|--" + o.lng;
    if (o.sh) return std::string("-") + o.sh;
    return std::string("--") + o.lng;
}

void take(Args& a, Opt& o, const std::string& val) {
    if (o.count >= o.maxc) {
        a.errors.push_back("excess option " + opt_name(o));
        return;
    }
    if (o.kind == Kind::Int) {
        long v;
        if (!parse_int(val, v)) {
            a.errors.push_back("invalid argument \"" + val + "\" to option " + opt_name(o));
            return;
        }
    } else if (o.kind == Kind::Dbl) {
        double v;
        if (!parse_dbl(val, v)) {
            a.errors.push_back("invalid argument \"" + val + "\" to option " + opt_name(o));
            return;
        }
    }
    o.vals.push_back(val);
    ++o.count;
}

void parse(Args& a, int argc, char** argv) {
    for (int i = 1; i < argc; ++i) {
        const std::string s = argv[i];
        if (s == "--") {
            for (++i; i < argc; ++i) a.errors.push_back(std::string("unexpected argument \"") + argv[i] + "\"");
            break;
        }
        if (s.size() > 2 && s[0] == '-' && s[1] == '-') {
            const size_t eq = s.find('=');
            const std::string name = s.substr(2, eq == std::string::npos ? std::string::npos : eq - 2);
            bool amb = false;
            Opt* o = a.by_long(name, amb);
            if (!o) {
                a.errors.push_back(std::string(amb ? "ambiguous" : "invalid") + " option \"" + s + "\"");
                continue;
            }
            if (o->kind == Kind::Lit) {
                if (eq != std::string::npos) a.errors.push_back("option --" + std::string(o->lng) + " takes no value");
                else if (o->count >= o->maxc) a.errors.push_back("excess option " + opt_name(*o));
                else ++o->count;
                continue;
            }
            if (eq != std::string::npos) take(a, *o, s.substr(eq + 1));
            else if (i + 1 < argc) take(a, *o, argv[++i]);
            else a.errors.push_back("option " + opt_name(*o) + " requires an argument");
            continue;
        }
        if (s.size() > 1 && s[0] == '-' && s[1] != '-') {
            // a cluster of short options; one that takes a value ends it
            for (size_t k = 1; k < s.size(); ++k) {
                Opt* o = a.by_short(s[k]);
                if (!o) {
                    a.errors.push_back(std::string("invalid option \"-") + s[k] + "\"");
                    break;
                }
                if (o->kind == Kind::Lit) {
                    if (o->count >= o->maxc) a.errors.push_back("excess option " + opt_name(*o));
                    else ++o->count;
                    continue;
                }
                if (k + 1 < s.size()) take(a, *o, s.substr(k + 1));
                else if (i + 1 < argc) take(a, *o, argv[++i]);
                else a.errors.push_back("option " + opt_name(*o) + " requires an argument");
                break;
            }
            continue;
        }
        a.errors.push_back("unexpected argument \"" + s + "\"");
    }
    for (Opt& o : a.opts)
        if (o.count < o.minc) a.errors.push_back("missing option " + opt_name(o));
}

void print_glossary(const Args& a) {
    for (const Opt& o : a.opts) {
        std::string syn;
        if (o.sh && o.lng) syn = std::string("-") + o.sh + ", --" + o.lng;
        else if (o.sh) syn = std::string("-") + o.sh;
        else syn = std::string("--") + o.lng;
        if (o.kind != Kind::Lit) syn += (o.lng ? "=" : " ") + std::string(o.datatype ? o.datatype : "<value>");
        std::printf("  %-45s %s\n", syn.c_str(), o.gloss);
    }
}

// --------------------------------------------------------------- config --

struct Config {
    bool compat = false;
    stt_geometry g{};
    stt_params p{};
    stt_of_params ofp{};
    int nframes = 1, nrounds = 1, views_per_round = 10;
    int outer = 20, pd_passes = 2, sart_sweeps = 1;
    std::string out = "Test/test", proj, lds = "lds.txt", init = "Input/rose.-%03d-frame-9iter.tiff",
                flow_dir = "FlowOutput";
    std::vector<float> sdg;  // --sdg values, default profile only
    bool batched = true, graph = true, dry_run = false, flow = true;
};

double dval(const Opt& o, size_t k, double dflt) { return k < o.vals.size() ? std::atof(o.vals[k].c_str()) : dflt; }
long ival(const Opt& o, size_t k, long dflt) {
    long v = dflt;
    if (k < o.vals.size()) parse_int(o.vals[k], v);
    return v;
}

int prior_of(const std::string& s, int& prior) {
    if (s == "ATV") prior = STT_PRIOR_ATV;
    else if (s == "TV") prior = STT_PRIOR_TV;
    else if (s == "STV") prior = STT_PRIOR_STV;
    else return 1;
    return 0;
}

std::string fmt(const char* f, ...) __attribute__((format(printf, 1, 2)));
std::string fmt(const char* f, ...) {
    char buf[2048];
    va_list ap;
    va_start(ap, f);
    std::vsnprintf(buf, sizeof buf, f, ap);
    va_end(ap);
    return buf;
}

// The path of frame f from a pattern holding exactly one integer conversion
// (%d, %i or %u with optional 0 flag and width, as "%03d"); "%%" is a
// literal percent sign.  Returns false for any other pattern.
bool frame_path(const std::string& pattern, int f, std::string& out) {
    out.clear();
    int conversions = 0;
    for (size_t i = 0; i < pattern.size(); ++i) {
        if (pattern[i] != '%') {
            out += pattern[i];
            continue;
        }
        if (i + 1 < pattern.size() && pattern[i + 1] == '%') {
            out += '%';
            ++i;
            continue;
        }
        size_t j = i + 1;
        const bool zero = j < pattern.size() && pattern[j] == '0';
        if (zero) ++j;
        int width = 0;
        while (j < pattern.size() && pattern[j] >= '0' && pattern[j] <= '9' && width < 100) width = width * 10 + (pattern[j++] - '0');
        if (j >= pattern.size() || (pattern[j] != 'd' && pattern[j] != 'i' && pattern[j] != 'u')) return false;
        std::string num = std::to_string(f);
        while ((int)num.size() < width) num.insert(num.begin(), zero ? '0' : ' ');
        out += num;
        ++conversions;
        i = j;
    }
    return conversions == 1;
}

int mkdirs(const std::string& dir) {
    if (dir.empty()) return 0;
    std::string cur;
    for (size_t i = 0; i <= dir.size(); ++i) {
        if (i == dir.size() || dir[i] == '/') {
            if (!cur.empty() && ::mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST) return -1;
        }
        if (i < dir.size()) cur += dir[i];
    }
    return 0;
}

std::string dir_of(const std::string& path) {
    const size_t s = path.rfind('/');
    return s == std::string::npos ? std::string() : path.substr(0, s);
}

int die(const std::string& msg) {
    std::fprintf(stderr, "stt-tomography: %s\n", msg.c_str());
    return 1;
}

int api(int rc, const char* what) {
    if (rc != STT_OK) std::fprintf(stderr, "stt-tomography: %s failed: %s\n", what, stt_last_error());
    return rc;
}
#define API(call)                                  \
    do {                                           \
        if (api((call), #call) != STT_OK) return 1; \
    } while (0)

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// ------------------------------------------------------------- writer ----

// Background writer of the run's dumps (warped frames, flows, volumes): a file
// is written while the GPU solves the next frame pair or PD pass, in the order
// it was queued.  Jobs own their samples; push() blocks while more than
// kMaxQueued bytes wait, so host memory stays bounded, and written buffers go
// back to a free list so the next dump reuses their pages.  The first failed
// write is reported by the next push() or by finish(); later jobs are dropped.
class Writer {
public:
    static constexpr size_t kMaxQueued = size_t(4) << 30;

    Writer() { th_ = std::thread([this] { loop(); }); }
    ~Writer() { finish(); }

    // a buffer of n floats, recycled when one of that size is free
    std::vector<float> take(size_t n) {
        std::lock_guard<std::mutex> lk(m_);
        for (size_t i = 0; i < free_.size(); ++i)
            if (free_[i].size() == n) {
                std::vector<float> v = std::move(free_[i]);
                free_.erase(free_.begin() + (long)i);
                free_bytes_ -= n * sizeof(float);
                return v;
            }
        return std::vector<float>(n);
    }

    // queue one file; false (error printed) once any write has failed
    bool push(bool tiff, std::string path, std::vector<float> data, const int dims[3], int c) {
        const size_t bytes = data.size() * sizeof(float);
        std::unique_lock<std::mutex> lk(m_);
        cv_.wait(lk, [&] { return failed_ || queued_ == 0 || queued_ + bytes <= kMaxQueued; });
        if (failed_) return report();
        q_.push_back(Job{tiff, std::move(path), std::move(data), {dims[0], dims[1], dims[2]}, c});
        queued_ += bytes;
        cv_.notify_all();
        return true;
    }

    // wait for every queued file; false (error printed) if a write failed
    bool finish() {
        {
            std::unique_lock<std::mutex> lk(m_);
            if (!th_.joinable()) return !failed_;
            stop_ = true;
            cv_.notify_all();
        }
        th_.join();
        std::lock_guard<std::mutex> lk(m_);
        return failed_ ? report() : true;
    }

private:
    struct Job {
        bool tiff;
        std::string path;
        std::vector<float> data;
        int d[3];
        int c;
    };

    bool report() {
        if (!reported_) std::fprintf(stderr, "stt-tomography: %s\n", error_.c_str());
        reported_ = true;
        return false;
    }

    void loop() {
        std::unique_lock<std::mutex> lk(m_);
        for (;;) {
            cv_.wait(lk, [&] { return stop_ || !q_.empty(); });
            if (q_.empty()) return;  // stop_ and drained
            Job j = std::move(q_.front());
            q_.pop_front();
            lk.unlock();
            int rc = STT_OK;
            std::string msg;
            if (!failed_) {
                rc = j.tiff ? stt_save_tiff(j.path.c_str(), j.data.data(), j.d[0], j.d[1], j.d[2], j.c)
                            : stt_save_analyze(j.path.c_str(), j.data.data(), j.d[0], j.d[1], j.d[2], j.c, nullptr);
                if (rc != STT_OK) msg = "writing " + j.path + " failed: " + stt_last_error();
            }
            lk.lock();
            if (rc != STT_OK && !failed_) {
                failed_ = true;
                error_ = msg;
            }
            const size_t bytes = j.data.size() * sizeof(float);
            queued_ -= bytes;
            if (free_bytes_ + bytes <= kMaxQueued) {
                free_bytes_ += bytes;
                free_.push_back(std::move(j.data));
            }
            cv_.notify_all();
        }
    }

    std::mutex m_;
    std::condition_variable cv_;
    std::deque<Job> q_;
    std::vector<std::vector<float>> free_;
    size_t queued_ = 0, free_bytes_ = 0;
    bool stop_ = false, failed_ = false, reported_ = false;
    std::string error_;
    std::thread th_;
};

// --------------------------------------------------------------- run ----

struct Plan {
    int W = 0, H = 0, Z = 0;
    std::vector<int> views;       // per frame
    std::vector<float> start;     // nframes x nrounds
};

int make_plan(const Config& c, Plan& pl) {
    stt_image_info info;
    if (stt_image_query(c.proj.c_str(), &info) != STT_OK) return die(stt_last_error());
    if (info.spectrum != 1) return die("the projection stack must have one channel");
    if (info.width != c.g.det[0] || info.height != c.g.det[1])
        return die(fmt("projection images are %dx%d, -u says %dx%d", info.width, info.height, c.g.det[0], c.g.det[1]));
    pl.W = info.width;
    pl.H = info.height;
    pl.Z = info.depth;
    pl.views.assign(c.nframes, 0);
    if (stt_split_sizes(pl.Z, c.nframes, pl.views.data()) != STT_OK) return die(stt_last_error());
    for (int f = 0; f < c.nframes; ++f)
        if (pl.views[f] != c.nrounds * c.views_per_round)
            return die(fmt("frame %d holds %d projections; %d rounds of %d views need %d", f, pl.views[f], c.nrounds,
                           c.views_per_round, c.nrounds * c.views_per_round));
    pl.start.assign((size_t)c.nframes * c.nrounds, 0.f);
    // --sdg is required by the flag table, so a run that wants lds.txt still
    // passes one value; only a full nframes x nrounds list replaces the file.
    if (!c.compat && c.sdg.size() == pl.start.size()) {
        pl.start = c.sdg;
    } else if (stt_read_start_angles(c.lds.c_str(), c.nframes, c.nrounds, pl.start.data()) != STT_OK) {
        return die(stt_last_error());
    }
    if (c.init != "zero") {
        for (int f = 0; f < c.nframes; ++f) {
            std::string path;
            if (!frame_path(c.init, f, path))
                return die("--init needs one integer conversion such as %03d, or 'zero'");
            if (stt_image_query(path.c_str(), &info) != STT_OK) return die(stt_last_error());
            if (info.width != c.g.vol[0] || info.height != c.g.vol[1] || info.depth != c.g.vol[2] || info.spectrum != 1)
                return die(fmt("'%s' is %dx%dx%dx%d, the volume is %dx%dx%d", path.c_str(), info.width, info.height,
                               info.depth, info.spectrum, c.g.vol[0], c.g.vol[1], c.g.vol[2]));
        }
    }
    return 0;
}

void print_plan(const Config& c, const Plan& pl) {
    std::printf("{\"profile\": \"%s\", \"vol\": [%d, %d, %d], \"spacing\": [%g, %g, %g], \"det\": [%d, %d], "
                "\"ds\": %g, \"sid\": %g, \"sdd\": %g, \"oxyz\": [%g, %g, %g], ",
                c.compat ? "reference" : "default", c.g.vol[0], c.g.vol[1], c.g.vol[2], c.g.spacing[0], c.g.spacing[1],
                c.g.spacing[2], c.g.det[0], c.g.det[1], c.g.ds, c.g.sid, c.g.sdd, c.g.oxyz[0], c.g.oxyz[1],
                c.g.oxyz[2]);
    std::printf("\"sigma\": %.9g, \"tau\": %.9g, \"lambda\": %.9g, \"k\": %.9g, \"e\": %.9g, \"huber\": %.9g, "
                "\"alpha\": %.9g, \"prior\": %d, \"bp\": %d, \"roi_cut\": [%d, %d, %d, %d, %d, %d, %d, %d], "
                "\"roi_jtop\": %d, ",
                c.p.sigma, c.p.tau, c.p.lambda_tv, c.p.k_tp, c.p.e_bc, c.p.huber, c.p.alpha, c.p.prior, c.p.bp,
                c.p.roi_cut[0], c.p.roi_cut[1], c.p.roi_cut[2], c.p.roi_cut[3], c.p.roi_cut[4], c.p.roi_cut[5],
                c.p.roi_cut[6], c.p.roi_cut[7], c.p.roi_jtop);
    std::printf("\"of_cut\": [%d, %d, %d, %d, %d, %d], \"nframes\": %d, \"nrounds\": %d, \"outer\": %d, "
                "\"pd_passes\": %d, \"sart_sweeps\": %d, \"flow\": %s, ",
                c.ofp.cut[0], c.ofp.cut[1], c.ofp.cut[2], c.ofp.cut[3], c.ofp.cut[4], c.ofp.cut[5], c.nframes,
                c.nrounds, c.outer, c.pd_passes, c.sart_sweeps, c.flow ? "true" : "false");
    std::printf("\"projections\": \"%s\", \"stack\": [%d, %d, %d], \"views\": [", c.proj.c_str(), pl.W, pl.H, pl.Z);
    for (size_t f = 0; f < pl.views.size(); ++f) std::printf("%s%d", f ? ", " : "", pl.views[f]);
    std::printf("], \"start_deg\": [");
    for (size_t k = 0; k < pl.start.size(); ++k) std::printf("%s%.9g", k ? ", " : "", pl.start[k]);
    std::printf("], \"init\": \"%s\", \"output\": \"%s\", \"flow_dir\": \"%s\"}\n", c.init.c_str(), c.out.c_str(),
                c.flow_dir.c_str());
}

int run(const Config& c, const Plan& pl) {
    const int F = c.nframes;
    const size_t V = (size_t)c.g.vol[0] * c.g.vol[1] * c.g.vol[2];
    const size_t D = (size_t)pl.W * pl.H;
    std::printf("%d %d %d %d %d\n", pl.W, pl.H, c.g.vol[0], c.g.vol[1], c.g.vol[2]);

    stt_ctx* ctx = nullptr;
    API(stt_ctx_create(&c.g, &c.p, F, pl.views.data(), &ctx));
    struct Guard {
        stt_ctx* c;
        ~Guard() { stt_ctx_destroy(c); }
    } guard{ctx};

    double t_io = now_s();
    // SetData: projections split on z into frames (ST_Tomography.cpp:355)
    {
        std::vector<float> stack((size_t)pl.Z * D);
        API(stt_load_image(c.proj.c_str(), stack.data(), stack.size()));
        size_t z0 = 0;
        for (int f = 0; f < F; ++f) {
            API(stt_set_projections(ctx, f, stack.data() + z0 * D));
            z0 += (size_t)pl.views[f];
        }
    }
    // InitViews: rounds of views_per_round views from each start angle (:887-934)
    for (int f = 0; f < F; ++f) {
        std::vector<float> q((size_t)4 * pl.views[f]);
        API(stt_view_rounds(pl.start.data() + (size_t)f * c.nrounds, c.nrounds, c.views_per_round, q.data()));
        API(stt_set_view_quaternions(ctx, f, q.data(), pl.views[f]));
    }
    // ReadStructuredFile: one volume per frame plus a copy of the last (VolumeData.cpp:115-126)
    {
        std::vector<float> vol(V, 0.f);
        for (int f = 0; f < F; ++f) {
            std::string path;
            if (c.init != "zero" && frame_path(c.init, f, path)) API(stt_load_image(path.c_str(), vol.data(), V));
            API(stt_set_volume(ctx, f, vol.data()));
        }
        API(stt_set_volume(ctx, F, vol.data()));
    }
    float gmax = 0.f;
    API(stt_ctx_get_gmax(ctx, &gmax));
    std::printf("G_Max %.9g\n", gmax);
    std::printf("inputs read and uploaded in %.3f s\n", now_s() - t_io);

    if (c.flow && F > 1 && mkdirs(c.flow_dir) != 0) return die("cannot create " + c.flow_dir);
    if (mkdirs(dir_of(c.out)) != 0) return die("cannot create " + dir_of(c.out));
    std::vector<int> idx_sart(F, 1);
    const int dims[3] = {c.g.vol[0], c.g.vol[1], c.g.vol[2]};
    Writer out;
    std::vector<int> its((size_t)c.ofp.levels * c.ofp.warps);
    for (int outer = 0; outer < c.outer; ++outer) {
        const double t_outer = now_s();
        double t_flow = 0.0, t_pd = 0.0, t_dump = 0.0;
        if (c.flow && F > 1) {
            std::printf("==============================================\n"
                        "Flow-Reconstruction-subOptimization starts\n"
                        "==============================================\n");
            for (int f = 0; f + 1 < F; ++f) {
                const double t0 = now_s();
                API(stt_optical_flow(ctx, f, &c.ofp, its.data()));
                const double t1 = now_s();
                t_flow += t1 - t0;
                std::vector<float> uvw = out.take(3 * V), warped = out.take(V);
                API(stt_get_buffer(ctx, STT_BUF_FLOW, f, uvw.data()));
                // warping::warp(dim, aabb, uvw, I2, -dt) on the device copies (:462)
                API(stt_warp_frame(ctx, f, -1.0f, STT_INTERP_CUBIC_CLAMPED, warped.data()));
                const std::string base = fmt("%s/%03d.%03d.%03d.frame%02d.iter%02d", c.flow_dir.c_str(), dims[0],
                                             dims[1], dims[2], f, outer);
                if (!out.push(true, base + ".tiff", std::move(warped), dims, 1)) return 1;
                if (!out.push(false, base + ".field.hdr", std::move(uvw), dims, 3)) return 1;
                t_dump += now_s() - t1;
                std::printf("flow pair %d: %.3f s, iterations per level/warp:", f, t1 - t0);
                for (int k : its) std::printf(" %d", k);
                std::printf("\n");
            }
        }
        if (outer == 0) {
            std::printf("outer iteration 0: %.3f s wall; flow %.3f s, warp + dump queueing %.3f s\n",
                        now_s() - t_outer, t_flow, t_dump);
            continue;
        }
        std::printf("==============================================\n"
                    "Image-Reconstruction-subOptimization starts\n"
                    "==============================================\n");
        // X_bar <- x before the first pass, duals at zero (ST_Tomography.cpp:494-500)
        if (outer == 1) API(stt_pd_init(ctx));
        for (int pass = 0; pass < c.pd_passes; ++pass) {
            const double t0 = now_s();
            if (c.sart_sweeps == 1) {
                API(stt_iterate(ctx, 1, c.batched ? 1 : 0, c.graph ? 1 : 0, nullptr));
            } else {
                for (int f = 0; f < F; ++f) {
                    API(stt_pd_dual(ctx, f));
                    API(stt_pd_primal(ctx, f));
                    for (int s = 0; s < c.sart_sweeps; ++s) API(stt_sart_sweep(ctx, f, nullptr, pl.views[f]));
                    API(stt_pd_extrapolate(ctx, f));
                }
            }
            API(stt_synchronize(ctx));
            const double t1 = now_s();
            t_pd += t1 - t0;
            std::printf("PD pass %d of outer iteration %d: %.3f s\n", pass, outer, t1 - t0);
            // ST_Tomography::SART's dump of each frame after its sweep (:786-805)
            for (int f = 0; f < F; ++f) {
                std::vector<float> vol = out.take(V);
                API(stt_get_volume(ctx, f, vol.data()));
                const std::string name = fmt("%s-%03d-frame-%diter.tiff", c.out.c_str(), f, idx_sart[f]);
                std::printf("%s\n", name.c_str());
                if (!out.push(true, name, std::move(vol), dims, 1)) return 1;
                ++idx_sart[f];
            }
            t_dump += now_s() - t1;
        }
        std::printf("outer iteration %d: %.3f s wall; flow %.3f s, PD %.3f s, warp + dump queueing %.3f s\n", outer,
                    now_s() - t_outer, t_flow, t_pd, t_dump);
    }
    // the files of the last iteration have no GPU work left to hide behind
    const double t_drain = now_s();
    if (!out.finish()) return 1;
    std::printf("remaining dumps written in %.3f s\n", now_s() - t_drain);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    Args a;
    Opt& sigma = a.add('s', "Sigma", Kind::Dbl, 0, 1, "<double>", "Sigma in volume density reconstruction (default as 0.2)");
    Opt& tau = a.add('t', "Tau", Kind::Dbl, 0, 1, "<double>", "Lambda in volume density reconstruction (default as 0.2)");
    Opt& lambda = a.add('l', "Lambda", Kind::Dbl, 0, 1, "<double>", "TV prior weight in volume density reconstruction ");
    Opt& tp = a.add('k', "TemporalPrior", Kind::Dbl, 0, 1, "<double>", "temporal prior weight in volume density reconstruction  ");
    Opt& bc = a.add('e', "Bright constancy prior", Kind::Dbl, 0, 1, "<double>", "Bright constancy weight in volume density reconstruction ");
    Opt& huber = a.add(0, "huber", Kind::Dbl, 0, 1, "<double>", "trade-off weight for huber norm epsilon ");
    Opt& alp = a.add('a', "Alpha", Kind::Dbl, 0, 1, "<double>", "stepsize for SART algorithm  (default as 0.3) ");
    Opt& outfile = a.add('o', nullptr, Kind::File, 0, 1, "<output>", "output 3d image (default as \"hello.tif\")");
    Opt& projfiles = a.add('i', "projsvolume", Kind::File, 1, 20, "<file>", "Input Dataset file(.mha.tif.tiff 2d bmp and other raw file are supported)");
    Opt& vols = a.add('f', "XYZ", Kind::Int, 1, 3, "<int>", "takes an integer value (defaults to 9)");
    Opt& imgwh = a.add('u', "imgUV", Kind::Int, 1, 2, "<int>", "w and h of projection image default:512 ");
    Opt& vs = a.add('c', "voxelspacing", Kind::Dbl, 0, 1, "<double>", "voxel spacing (default as 1)");
    Opt& ds = a.add('d', "detector spacing", Kind::Dbl, 0, 1, "<double>", "detector spacing (default as 1)");
    Opt& sid = a.add('g', "sid", Kind::Dbl, 0, 1, "<double>", " Source to Iso-Object Distance  (default as 600mm)");
    Opt& sdd = a.add('j', "sdd", Kind::Dbl, 0, 1, "<double>", " Source to Detector Distance (default as 1000mm)");
    Opt& oxyz = a.add(0, "oxyz", Kind::Dbl, 1, 3, "<double>", " offset x y z of the center of volume (default as 0)");
    Opt& sdg = a.add(0, "sdg", Kind::Dbl, 1, 20, "<double>", " start degree for each proj sequence");
    Opt& nframe = a.add('v', "nframes", Kind::Int, 0, 1, "<int>", "Number of frames");
    Opt& nrounds = a.add('r', "nframes", Kind::Int, 0, 1, "<int>", "Number of rounds for each frames");
    Opt& algo = a.add('b', "AlgorithmIter", Kind::Int, 0, 1, "<int>", "Number of algorithm iterations, default as 20");
    Opt& sart = a.add('p', "SartIter", Kind::Int, 0, 1, "<int>", "Number of SART nested iterations, default as 1");
    Opt& prior = a.add(0, "prior", Kind::Str, 0, 1, "{STV,SAD,ATV}", "specify the prior you are applying from {STV,SAD,TV} , default as TV");
    Opt& bp = a.add(0, "bp", Kind::Str, 0, 1, "{Voxelbased,Raybased}", "specify the backprojection methods you are applying from {Voxelbased,Raybased}, default as voxelbased");
    Opt& help = a.add('h', "help", Kind::Lit, 0, 1, nullptr, "print this help and exit");
    Opt& version = a.add(0, "version", Kind::Lit, 0, 1, nullptr, "print version information and exit");
    // options of this driver (no counterpart in main.cpp)
    Opt& compat = a.add(0, "compat", Kind::Str, 0, 1, "{reference}", "run with the values the reference actually uses (see the source header)");
    Opt& lds = a.add(0, "lds", Kind::File, 0, 1, "<file>", "start-angle file (default lds.txt)");
    Opt& init = a.add(0, "init", Kind::Str, 0, 1, "<pattern|zero>", "initial volume per frame, printf pattern (default Input/rose.-%03d-frame-9iter.tiff)");
    Opt& flowdir = a.add(0, "flow-dir", Kind::File, 0, 1, "<dir>", "directory of the flow outputs (default FlowOutput)");
    Opt& joint = a.add(0, "joint-iters", Kind::Int, 0, 1, "<int>", "outer flow/reconstruction iterations (default 20)");
    Opt& noflow = a.add(0, "no-flow", Kind::Lit, 0, 1, nullptr, "skip the flow sub-optimisation (flows stay zero)");
    Opt& perframe = a.add(0, "per-frame", Kind::Lit, 0, 1, nullptr, "one frame per launch instead of the batched form (same results)");
    Opt& nograph = a.add(0, "no-graph", Kind::Lit, 0, 1, nullptr, "issue PD passes without hipGraph replay");
    Opt& dry = a.add(0, "dry-run", Kind::Lit, 0, 1, nullptr, "resolve flags and inputs, print them as JSON, touch no device");

    parse(a, argc, argv);
    const char* progname = "stt-tomography";
    if (help.count > 0) {
        std::printf("Usage: %s [options]\n", progname);
        print_glossary(a);
        return 0;
    }
    if (version.count > 0) {
        std::printf("%s: MI355X space-time tomography, C-ABI version %d\n", progname, stt_version());
        return 0;
    }
    if (!a.errors.empty()) {
        for (const std::string& e : a.errors) std::printf("%s: %s\n", progname, e.c_str());
        std::printf("Try '%s --help' for more information.\n", progname);
        return 1;
    }

    Config c;
    c.compat = compat.count > 0;
    if (c.compat && compat.vals[0] != "reference") return die("--compat takes only 'reference'");
    if (c.compat) stt_params_reference(&c.p);
    else stt_params_default(&c.p);
    stt_of_params_reference(&c.ofp);
    if (!c.compat)
        for (int k = 0; k < 6; ++k) c.ofp.cut[k] = 0;  // whole grid

    // main.cpp:67-95 defaults, then the flags
    int vol[3] = {63, 63, 63};
    for (int k = 0; k < 3; ++k) vol[k] = (int)ival(vols, k, vol[k]);
    int det[2] = {512, 512};
    for (int k = 0; k < 2; ++k) det[k] = (int)ival(imgwh, k, det[k]);
    const double spacing = dval(vs, 0, 1.0);
    for (int k = 0; k < 3; ++k) {
        c.g.vol[k] = vol[k];
        c.g.spacing[k] = (float)spacing;
        c.g.oxyz[k] = dval(oxyz, k, 0.0);
    }
    c.g.det[0] = det[0];
    c.g.det[1] = det[1];
    c.g.ds = dval(ds, 0, 1.0);
    c.g.sid = dval(sid, 0, 400.0);
    c.g.sdd = dval(sdd, 0, 200.0);
    c.p.lambda_tv = dval(lambda, 0, 0.1);
    c.p.k_tp = dval(tp, 0, 0.0);
    c.p.e_bc = dval(bc, 0, 0.0);
    c.p.huber = dval(huber, 0, 0.1);
    c.p.alpha = dval(alp, 0, 0.1);
    const std::string bps = bp.count ? bp.vals[0] : "Voxelbased";
    // ST_Tomography.cpp:221-226: anything but "Raybased" is voxel-based
    c.p.bp = bps == "Raybased" ? STT_BP_RAY : STT_BP_VOXEL;
    c.nframes = (int)ival(nframe, 0, 1);
    c.nrounds = (int)ival(nrounds, 0, 1);
    c.out = outfile.count ? outfile.vals[0] : "Test/test";
    c.proj = projfiles.vals.empty() ? "proj" : projfiles.vals[0];  // only the first is read (:355)
    if (lds.count) c.lds = lds.vals[0];
    if (init.count) c.init = init.vals[0];
    if (flowdir.count) c.flow_dir = flowdir.vals[0];
    c.flow = noflow.count == 0;
    c.batched = perframe.count == 0;
    c.graph = nograph.count == 0;
    c.dry_run = dry.count > 0;
    if (c.compat) {
        c.outer = 20;
        c.pd_passes = 2;
        c.sart_sweeps = 1;
    } else {
        c.p.sigma = dval(sigma, 0, 0.3);
        c.p.tau = dval(tau, 0, 0.3);
        if (prior.count && prior_of(prior.vals[0], c.p.prior)) return die("--prior must be ATV, TV or STV");
        c.pd_passes = (int)ival(algo, 0, 2);
        c.sart_sweeps = (int)ival(sart, 0, 2);
        for (const std::string& s : sdg.vals) c.sdg.push_back((float)std::atof(s.c_str()));
    }
    if (joint.count) c.outer = (int)ival(joint, 0, 20);
    if (c.nframes < 1 || c.nrounds < 1) return die("-v and -r must be at least 1");
    if (c.outer < 0 || c.pd_passes < 0 || c.sart_sweeps < 1) return die("iteration counts must be positive");

    Plan pl;
    if (make_plan(c, pl)) return 1;
    if (c.dry_run) {
        print_plan(c, pl);
        return 0;
    }
    const double t0 = now_s();
    const int rc = run(c, pl);
    if (rc == 0) std::printf("done in %.3f s\n", now_s() - t0);
    return rc;
}
