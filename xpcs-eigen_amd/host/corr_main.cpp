// `corr` -- drop-in replacement of the reference's command line
// (src/xpcs/main.cpp:86-479): multi-tau, or two-time when <entry>/analysis_type
// is "Twotime", running the correlator on the GPU through the C ABI.
//
//   corr <config.hdf5> [data.imm] [--imm=PATH] [--entry=/xpcs] [--exchange=GROUP]
//        [--inpath=P --outpath=Q] [--g2out] [--darkout] [--frame_threading]
//        [--ufxc | --rigaku | --hdf5 [--transposed]] [--quiet]
//
// Flags use the gflags syntax of the reference (--name=value, --name value,
// --bool / --nobool).  The IMM path comes from --imm, else the optional second
// positional argument (as README.md documents), else <entry>/input_file_local,
// with the --inpath prefix replaced by --outpath.  Results are written into the
// configuration file under <entry>/output_data (or --exchange) with the
// reference's dataset names, shapes and dtypes.  --frame_threading selects the
// reference's other two-time loop order (corr.cpp:574-779); both compute the
// same C2T, so it changes nothing here.
//
// Input: IMM, or with --ufxc / --rigaku / --hdf5 (the reference's flags, same
// precedence) UFXC or Rigaku packed words, or HDF5 frames (/entry/data/data);
// --transposed is accepted (the reference's HDF5 reader does not use it).
//
// normalize_by_framesum = 1 scales every value by mean(frameSum) / frameSum of
// its frame before the correlation (FP64 sums, main.cpp:313-336).
//
// Exit status: 0 ok, 1 usage or configuration error (including an unknown
// two-time smoothing method), 2 option outside this build (--frameout),
// 3 unreadable or malformed input file, 4 GPU error, 5 cannot write results.
// (The reference exits 0 on most failures.)

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <chrono>
#include <cstdarg>
#include <cstdlib>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "config.h"
#include "h5lite.h"
#include "xpcs_mt.h"

namespace {

bool g_quiet = false;

void info(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
void info(const char* fmt, ...)
{
    if (g_quiet) return;
    va_list ap;
    va_start(ap, fmt);
    std::printf("[info] ");
    std::vprintf(fmt, ap);
    std::printf("\n");
    std::fflush(stdout);
    va_end(ap);
}

// RAII scope timer printing "<name> took <t> ms" (the reference's Benchmark,
// src/xpcs/benchmark.h:56-86).
struct Scope {
    std::string name;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    explicit Scope(std::string n) : name(std::move(n)) {}
    ~Scope()
    {
        double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        info("%s took %.3f ms", name.c_str(), ms);
    }
};

int die(int code, const std::string& msg)
{
    std::fprintf(stderr, "corr: %s\n", msg.c_str());
    return code;
}

struct Flags {
    std::map<std::string, std::string> kv;
    std::vector<std::string> pos;
    std::string bad;
};

const char* kBoolFlags[] = {"g2out", "darkout", "ufxc", "rigaku", "hdf5", "frame_threading", "transposed", "quiet"};
const char* kValueFlags[] = {"imm", "inpath", "outpath", "exchange", "entry", "frameout"};

bool is_in(const char* const* set, size_t n, const std::string& s)
{
    for (size_t i = 0; i < n; i++)
        if (s == set[i]) return true;
    return false;
}

Flags parse(int argc, char** argv)
{
    Flags f;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        if (a.size() < 2 || a[0] != '-') {
            f.pos.push_back(a);
            continue;
        }
        std::string body = a.substr(a[1] == '-' ? 2 : 1);
        std::string name = body, val;
        const size_t eq = body.find('=');
        bool has_val = eq != std::string::npos;
        if (has_val) {
            name = body.substr(0, eq);
            val = body.substr(eq + 1);
        }
        if (is_in(kBoolFlags, sizeof(kBoolFlags) / sizeof(*kBoolFlags), name)) {
            f.kv[name] = has_val ? val : "true";
        } else if (name.rfind("no", 0) == 0 && is_in(kBoolFlags, sizeof(kBoolFlags) / sizeof(*kBoolFlags), name.substr(2))) {
            f.kv[name.substr(2)] = "false";
        } else if (is_in(kValueFlags, sizeof(kValueFlags) / sizeof(*kValueFlags), name)) {
            if (!has_val) {
                if (i + 1 >= argc) {
                    f.bad = "flag --" + name + " needs a value";
                    return f;
                }
                val = argv[++i];
            }
            f.kv[name] = val;
        } else {
            f.bad = "unknown flag " + a;
            return f;
        }
    }
    return f;
}

bool flag_true(const Flags& f, const char* name)
{
    auto it = f.kv.find(name);
    return it != f.kv.end() && (it->second == "true" || it->second == "1" || it->second == "yes");
}

std::string flag_str(const Flags& f, const char* name)
{
    auto it = f.kv.find(name);
    return it == f.kv.end() ? std::string() : it->second;
}

struct Mapped {
    void* p = nullptr;
    size_t n = 0;
    ~Mapped()
    {
        if (p && p != MAP_FAILED) munmap(p, n);
    }
};

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 2) return die(1, "usage: corr <config.hdf5> [data.imm] [--imm=PATH] [--entry=/xpcs] [--exchange=GROUP] [--g2out] [--darkout]");
    Flags fl = parse(argc, argv);
    if (!fl.bad.empty()) return die(1, fl.bad);
    if (fl.pos.empty()) return die(1, "missing configuration file argument");
    g_quiet = flag_true(fl, "quiet");
    Scope total("Total");
    // input format (main.cpp:204-220: --ufxc, else --rigaku, else --hdf5, else IMM)
    enum { kImm, kUfxc, kRigaku, kHdf5 } fmt = kImm;
    if (flag_true(fl, "ufxc")) fmt = kUfxc;
    else if (flag_true(fl, "rigaku")) fmt = kRigaku;
    else if (flag_true(fl, "hdf5")) fmt = kHdf5;
    if (!flag_str(fl, "frameout").empty() && std::atoi(flag_str(fl, "frameout").c_str()) > 0)
        return die(2, "--frameout (debug frame dump) is not built in this tool");

    std::string entry = flag_str(fl, "entry").empty() ? "/xpcs" : flag_str(fl, "entry");
    info("H5 metadata path %s", entry.c_str());
    std::string why;
    if (!xpcs::h5::available(&why)) return die(1, "HDF5 unavailable: " + why);
    xpcs::JobConfig cfg;
    {
        Scope s("Configuration Total");
        if (!xpcs::load_config(fl.pos[0], entry, &cfg, &why)) return die(1, why);
    }
    // main.cpp:161-164: two-time needs a known smoothing method
    if (cfg.twotime && cfg.smoothing == 0) return die(1, "Smoothing method is not valid");
    if (cfg.twotime && cfg.twotime_window < 1) return die(1, "twotime2onetime_window_size must be >= 1");
    if (!flag_str(fl, "imm").empty()) cfg.imm_path = flag_str(fl, "imm");
    else if (fl.pos.size() > 1) cfg.imm_path = fl.pos[1];
    const std::string inp = flag_str(fl, "inpath"), outp = flag_str(fl, "outpath");
    if (!inp.empty() && !outp.empty()) {
        const size_t pos = cfg.imm_path.find(inp);
        if (pos != std::string::npos) cfg.imm_path.replace(pos, inp.size(), outp);
    }
    if (!flag_str(fl, "exchange").empty()) cfg.output_group = flag_str(fl, "exchange");
    info("Processing IMM file at path %s..", cfg.imm_path.c_str());

    // map the IMM / UFXC / Rigaku file, or read the HDF5 frames
    Mapped imm;
    std::vector<uint16_t> h5_16;
    std::vector<uint32_t> h5_32;
    int64_t h5_frames = 0;
    if (fmt == kHdf5) {
        // Hdf5::Hdf5 / NextFrames (src/xpcs/io/hdf5.cpp:62-199): /entry/data/data,
        // frames x (height x width) unsigned 16- or 32-bit values
        Scope s("Reading HDF5 frames");
        xpcs::h5::File in;
        if (!in.open(cfg.imm_path, false, &why)) return die(3, "cannot open HDF5 file " + cfg.imm_path + ": " + why);
        std::vector<xpcs::h5::hsize_t> dd;
        if (!in.read_frames("/entry/data/data", &h5_16, &h5_32, &dd) || dd.size() != 3)
            return die(3, "cannot read /entry/data/data (3-D, 16- or 32-bit unsigned) from " + cfg.imm_path);
        if ((int64_t)dd[1] * (int64_t)dd[2] != (int64_t)cfg.width * cfg.height)
            return die(3, "HDF5 frames do not hold x_dimension*y_dimension values");
        h5_frames = (int64_t)dd[0];
    } else {
        int fd = open(cfg.imm_path.c_str(), O_RDONLY);
        if (fd < 0) return die(3, "cannot open IMM file " + cfg.imm_path);
        struct stat st;
        fstat(fd, &st);
        imm.n = (size_t)st.st_size;
        imm.p = imm.n ? mmap(nullptr, imm.n, PROT_READ, MAP_PRIVATE, fd, 0) : nullptr;
        close(fd);
        if (!imm.p || imm.p == MAP_FAILED) return die(3, "cannot map IMM file " + cfg.imm_path);
        info("File size %.3f MB", imm.n / 1048576.0);
    }
    // UFXC, Rigaku and HDF5 readers report compressed data: no dark frames (main.cpp:227-239)
    if (fmt != kImm) cfg.dark_frames = 0;

    const int32_t F = cfg.frames();
    info("Data frames=%d stride=%lld average=%lld", F, (long long)cfg.stride, (long long)cfg.average);
    xpcs_mt_desc d{};
    d.width = cfg.width;
    d.height = cfg.height;
    d.frames = F;
    d.delays_per_level = cfg.delays_per_level;
    d.stride = (int32_t)cfg.stride;
    d.average = (int32_t)cfg.average;
    d.frame_from = cfg.frame_from();
    d.dark_frames = cfg.dark_frames;
    d.lld = cfg.lld;
    d.sigma = cfg.sigma;
    d.dqmap = cfg.dqmap.data();
    d.sqmap = cfg.sqmap.data();
    d.flatfield = cfg.flatfield_enabled ? cfg.flatfield.data() : nullptr;
    d.flags = XPCS_MT_SIDE_STATS | (cfg.normalize_by_framesum ? XPCS_MT_NORMALIZE_BY_FRAMESUM : 0u);
    xpcs_mt_plan* plan = nullptr;
    int st = xpcs_mt_plan_create(&d, &plan);
    if (st == XPCS_ERR_UNSUPPORTED) return die(2, xpcs_last_error());
    if (st) return die(st == XPCS_ERR_NO_DEVICE || st == XPCS_ERR_HIP || st == XPCS_ERR_NOMEM ? 4 : 1, xpcs_last_error());
    struct PlanGuard {
        xpcs_mt_plan* p;
        ~PlanGuard() { xpcs_mt_plan_destroy(p); }
    } guard{plan};
    xpcs_mt_info pi{};
    xpcs_mt_plan_info(plan, &pi);
    const int T = pi.taus, Q = pi.dynamic_bins, S = pi.static_bins;

    {
        Scope s("Loading data");
        if (fmt == kUfxc) st = xpcs_mt_ingest_packed(plan, imm.p, imm.n, XPCS_FMT_UFXC);
        else if (fmt == kRigaku) st = xpcs_mt_ingest_packed(plan, imm.p, imm.n, XPCS_FMT_RIGAKU);
        else if (fmt == kHdf5)
            st = h5_32.empty() ? xpcs_mt_ingest_frames(plan, h5_16.data(), h5_frames, 2)
                               : xpcs_mt_ingest_frames(plan, h5_32.data(), h5_frames, 4);
        else st = xpcs_mt_ingest_imm(plan, imm.p, imm.n);
        if (st) return die(st == XPCS_ERR_UNSUPPORTED ? 2 : 3, xpcs_last_error());
    }
    {
        Scope s("Computing G2 MultiTau + Normalizing Data");
        st = xpcs_mt_execute(plan, nullptr);
        if (!st) st = xpcs_mt_synchronize(plan);
        if (st) return die(st == XPCS_ERR_RANGE ? 3 : 4, xpcs_last_error());
    }
    std::vector<double> g2((size_t)T * Q), se((size_t)T * Q);
    std::vector<float> pix((size_t)cfg.width * cfg.height), fsum((size_t)2 * F), pmean((size_t)S);
    const int nwin = cfg.static_window > 0 ? F / cfg.static_window : 0;
    std::vector<float> ppart((size_t)std::max(nwin, 0) * S);
    st = xpcs_mt_read_g2(plan, g2.data(), se.data());
    if (!st)
        st = xpcs_mt_read_side(plan, cfg.static_window, pix.data(), fsum.data(), pmean.data(),
                               nwin > 0 ? ppart.data() : nullptr);
    if (st) return die(4, xpcs_last_error());
    std::vector<int32_t> lev((size_t)T), tau((size_t)T);
    xpcs_mt_schedule(plan, lev.data(), tau.data());

    // two-time (Corr::twotime, corr.cpp:562-924): C2T per listed q-bin,
    // g2full, g2partials and the smoothing image sg
    int32_t tt_nq = 0, tt_rows = 0, tt_cols = 0, tt_total = 0;
    std::vector<int32_t> tt_q;
    if (cfg.twotime) {
        Scope s("Computing G2 TwoTimes");
        xpcs_tt_desc td{};
        td.qbins = cfg.qphi_bins.data();
        td.nq = (int32_t)cfg.qphi_bins.size();
        td.smoothing = cfg.smoothing;
        td.average = cfg.smoothing_average ? 1 : 0;
        td.window = cfg.twotime_window;
        st = xpcs_mt_twotime(plan, &td, nullptr);
        tt_q.resize(std::max<size_t>(cfg.qphi_bins.size(), 1));
        if (!st)
            st = xpcs_mt_twotime_info(plan, &tt_nq, tt_q.data(), nullptr, (int32_t)tt_q.size(), &tt_rows, &tt_cols,
                                      &tt_total);
        if (st) return die(st == XPCS_ERR_RANGE ? 3 : 4, xpcs_last_error());
    }

    // timestamps of the data frames (row 0: 1-based index, row 1: header value;
    // the other readers give frame indices: UFXC the raw frame from 0 (its
    // SkipFrames does nothing), HDF5 the raw frame after the skip, Rigaku the
    // output frame (and nothing past `frames`))
    const int32_t FR = cfg.real_frames();
    std::vector<double> clk((size_t)2 * FR, 0.0), tck((size_t)2 * FR, 0.0);
    if (fmt != kImm) {
        for (int32_t i = 0; i < FR; i++) {
            clk[(size_t)i] = tck[(size_t)i] = (double)(i + 1);
            double v = fmt == kUfxc ? (double)i : fmt == kHdf5 ? (double)(cfg.frame_from() + i) : (double)i;
            if (fmt == kRigaku && i >= F) v = 0.0;
            clk[(size_t)FR + i] = tck[(size_t)FR + i] = v;
        }
    } else {
        int64_t n = 0;
        int32_t comp = 0;
        xpcs_imm_index(imm.p, imm.n, nullptr, nullptr, nullptr, nullptr, 0, &n, &comp);
        std::vector<double> el((size_t)n);
        std::vector<int32_t> ct((size_t)n);
        xpcs_imm_index(imm.p, imm.n, nullptr, nullptr, el.data(), ct.data(), n, &n, &comp);
        int64_t first = (!comp && cfg.dark_frames > 0) ? cfg.dark_frames : 0;
        if (cfg.frame_from() > first) first = cfg.frame_from();
        for (int32_t i = 0; i < FR; i++) {
            clk[(size_t)i] = tck[(size_t)i] = (double)(i + 1);
            if (first + i < n) {
                clk[(size_t)FR + i] = el[(size_t)(first + i)];
                tck[(size_t)FR + i] = (double)ct[(size_t)(first + i)];
            }
        }
    }

    {
        Scope s("Writing results");
        xpcs::h5::File out;
        if (!out.open(cfg.path, false, &why)) return die(5, why);
        std::string G = cfg.output_group;
        if (G.empty() || G[0] != '/') G = "/" + G;
        auto P = [&](const char* n) { return G + "/" + n; };
        using V = std::vector<xpcs::h5::hsize_t>;
        std::vector<float> g2f(g2.size()), sef(se.size()), tauf((size_t)T);
        for (size_t i = 0; i < g2.size(); i++) {
            g2f[i] = (float)g2[i];
            sef[i] = (float)se[i];
        }
        for (int i = 0; i < T; i++) tauf[(size_t)i] = (float)tau[(size_t)i];
        bool ok = out.write_f32(P("pixelSum"), pix.data(), V{(unsigned long long)cfg.height, (unsigned long long)cfg.width}) &&
                  out.write_f32(P("frameSum"), fsum.data(), V{2, (unsigned long long)F}) &&
                  out.write_f32(P("partition-mean-total"), pmean.data(), V{1, (unsigned long long)S}) &&
                  out.write_f32(P("partition-mean-partial"), ppart.data(), V{(unsigned long long)std::max(nwin, 0), (unsigned long long)S}) &&
                  out.write_f32(P("partition_norm_factor"), &cfg.norm_factor, V{1, 1}) &&
                  out.write_f64(P("timestamp_clock"), clk.data(), V{2, (unsigned long long)FR}) &&
                  out.write_f64(P("timestamp_tick"), tck.data(), V{2, (unsigned long long)FR});
        if (ok && !cfg.twotime)
            ok = out.write_f32(P("tau"), tauf.data(), V{1, (unsigned long long)T}) &&
                 out.write_f32(P("norm-0-g2"), g2f.data(), V{(unsigned long long)T, (unsigned long long)Q}) &&
                 out.write_f32(P("norm-0-stderr"), sef.data(), V{(unsigned long long)T, (unsigned long long)Q});
        if (ok && cfg.twotime) {
            // C2T_all/g2_%05d [frames][frames] (chunked, deflate 6), g2full
            // [frames][q], g2partials [window][partials][q], sg [rows][cols]
            const size_t FF = (size_t)F * (size_t)F;
            std::vector<double> c2t(FF);
            std::vector<float> c2tf(FF);
            for (int32_t i = 0; ok && i < tt_nq; i++) {
                st = xpcs_mt_read_twotime(plan, i, c2t.data());
                if (st) return die(4, xpcs_last_error());
                for (size_t k = 0; k < FF; k++) c2tf[k] = (float)c2t[k];
                char name[32];
                std::snprintf(name, sizeof(name), "C2T_all/g2_%05d", tt_q[(size_t)i]);
                ok = out.write_f32_deflate(P(name), c2tf.data(), V{(unsigned long long)F, (unsigned long long)F}, 6);
            }
            const size_t W = (size_t)cfg.twotime_window, tot = (size_t)tt_total, nq = (size_t)tt_nq;
            std::vector<double> full((size_t)F * nq), part(W * tot * nq), sg((size_t)tt_rows * tt_cols);
            if (ok) {
                st = xpcs_mt_read_twotime_1d(plan, full.data(), part.data(), sg.data());
                if (st) return die(4, xpcs_last_error());
                auto f32 = [](const std::vector<double>& v) { return std::vector<float>(v.begin(), v.end()); };
                const std::vector<float> fullf = f32(full), partf = f32(part), sgf = f32(sg);
                ok = out.write_f32(P("g2full"), fullf.data(), V{(unsigned long long)F, nq}) &&
                     out.write_f32(P("g2partials"), partf.data(), V{W, tot, nq}) &&
                     out.write_f32(P("sg"), sgf.data(), V{(unsigned long long)tt_rows, (unsigned long long)tt_cols});
            }
        }
        if (ok && flag_true(fl, "g2out") && !cfg.twotime) {
            const size_t n = (size_t)cfg.width * cfg.height * T;
            std::vector<float> a(n), b(n), c(n);
            st = xpcs_mt_read_ref_f32(plan, a.data(), b.data(), c.data());
            if (st) return die(4, xpcs_last_error());
            V dd{(unsigned long long)T, (unsigned long long)cfg.width * cfg.height};
            ok = out.write_f32(P("G2"), a.data(), dd) && out.write_f32(P("IP"), b.data(), dd) &&
                 out.write_f32(P("IF"), c.data(), dd);
        }
        if (ok && flag_true(fl, "darkout") && cfg.dark_frames > 0) {
            std::vector<double> da((size_t)cfg.width * cfg.height), ds((size_t)cfg.width * cfg.height);
            st = xpcs_mt_read_dark(plan, da.data(), ds.data());
            if (!st) {
                V dd{(unsigned long long)cfg.height, (unsigned long long)cfg.width};
                ok = out.write_f64(P("DarkAvg"), da.data(), dd) && out.write_f64(P("DarkStd"), ds.data(), dd);
            }
        }
        if (!ok) return die(5, "cannot write results into " + cfg.path + " group " + G);
    }
    info("pixels=%d taus=%d q=%d events=%lld", pi.pixels, T, Q, (long long)pi.events);
    return 0;
}
