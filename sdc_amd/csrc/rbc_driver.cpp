// rbc_driver.cpp -- the simulation driver of libsdc_rbc: the reference's main program
// (gpu_flow/code/main.cu) on top of the C-ABI.  rbc_run_main takes the reference's argument
// set (main.cu:61-229, including the names README.txt:20-66 still lists: -delta_t, -final_t),
// sets up a context, runs iCeil(t / dt) steps (main.cu:279, util/util.cu:13-17) with the
// mean-flow forcing before each step (-mean_F, main.cu:567-568), advances the tracers
// (main.cu:628-660) and writes the state files every (steps / nfiles) steps as
// theta_%08d.bin ... G_%08d.bin (main.cu:697-728), the tracks as particles_le%.0le_%08d.trkss
// and the tracer second moments as secondmoment_%d.dat.  The sdc_amd/rbc_run executable is a
// main() around it.
//
// Differences from the reference, on purpose:
//   - the precision is a run-time option (-precision f32 | f64 | bf16x3; f32 is the reference's
//     default build, f64 its PRECISION_DOUBLE build); real arguments are parsed in that build's
//     real type (float for f32 and bf16x3) as CUDA_FLOAT_REAL is;
//   - the working directory is a path prefix instead of a chdir (relative -files paths are
//     resolved against it, as after the reference's chdir);
//   - a step that ends non-finite stops the run with an error (the reference carries NaN on);
//   - -write_history also writes the AB2 history next to each state file, -history_files
//     restores one, so a restart continues with AB2 instead of an Euler step;
//   - -tracer_mode spectral selects the spectral RK2 tracer (default: the reference's RK4).
// -freeze_circle n and -temp_circle n run the freeze and heat-source operators before every step
// from step n on (main.cu:569-590).  -penalization rectangle|circle|diskarray|everywhere (an option
// of this driver) runs the Brinkman terms the PENALIZED builds compile in, with -eta.  The pattern starts (-init_rectangle, -init_dislocation,
// -init_isr, -init_sv, -init_cr, -init_giantspiral with -init_mode) go through rbc_init_pattern.
// -port (USE_SERVER)
// and -snapshot_freq (WRITE_SNAPSHOT) are accepted and ignored, as in the reference's default
// build.
#include <hip/hip_runtime.h>
#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "rbc_abi.h"

namespace {

// ---- argument lookup with the reference parser's rules (util/inputparser.cpp) ------------
//   values: the first occurrence of the key, `count` values after it; a value that does not
//   parse is an error; strings: the last occurrence; a missing mandatory key is an error.
struct Args {
    std::vector<std::string> v;
    std::string err;
    int pos(const char* name) const {
        for (size_t i = 0; i < v.size(); ++i)
            if (v[i] == name) return (int)i;
        return -1;
    }
    bool opt(const char* name) const { return pos(name) >= 0; }
    template <typename T>
    bool get(std::vector<T>& out, int count, const char* name, bool mandatory) {
        int i = pos(name);
        bool found = i >= 0;
        if (found && count > (int)v.size() - (i + 1)) found = false;
        if (found) {
            out.resize(count);
            for (int k = 0; k < count; ++k) {
                std::stringstream s(v[i + 1 + k]);
                T val;
                s >> val;
                if (s.fail()) {
                    err = std::string("Parsing argument failed: ") + name;
                    return false;
                }
                out[k] = val;
            }
        }
        if (mandatory && !found) {
            err = std::string("Mandatory argument not found: ") + name;
            return false;
        }
        return true;
    }
    template <typename T>
    bool get1(T& out, const char* name, bool mandatory) {
        std::vector<T> o;
        if (!get(o, 1, name, mandatory)) return false;
        if (!o.empty()) out = o[0];
        return true;
    }
    bool strings(std::vector<std::string>& out, int count, const char* name) const {
        bool found = false;
        for (int i = 0; i + count < (int)v.size(); ++i)
            if (v[i] == name) {
                found = true;
                out.assign(v.begin() + i + 1, v.begin() + i + 1 + count);
            }
        return found;
    }
};

int iceil(double x) {  // util/util.cu:13-17
    const int ix = (int)x;
    return x > (double)ix ? ix + 1 : ix;
}

// ---- the .trkss track format (output/track_movie.cpp Write, track_writer.cu) -------------
// "TRKSS", version 'a', uint32 tracks, uint32 fps, uint32 length + description, uint32 first
// and last frame time, uint32 frames; per frame uint32 time and float x, y, z per track.
struct TrackWriter {
    std::string description;
    int period = 1, time = 0, offset = 0;
    uint32_t ntracks = 0;
    std::vector<uint32_t> times;
    std::vector<float> xyz;  // per frame: x, y, z of every track, interleaved per track
    // ≙ track_writer::append: frames [slot][particle] of one read-out of the tracer ring.  As in
    // the reference (track_writer.cu:45-61), offset changes only when a frame is taken: a read-out
    // of fewer slots than the offset leaves it as it was.
    void append(const float* x, const float* y, const float* z, int np, int nframes) {
        if (times.empty()) ntracks = (uint32_t)np;
        for (int t = offset; t < nframes; t += period) {
            times.push_back((uint32_t)time);
            for (int p = 0; p < np; ++p) {
                const size_t e = (size_t)t * np + p;
                xyz.push_back(x[e]);
                xyz.push_back(y[e]);
                xyz.push_back(z[e]);
            }
            time += period;
            offset = t - nframes + period;
        }
    }
    // ≙ track_writer::write_file: writes and clears the frames (time keeps counting)
    bool write(const std::string& path) {
        if (times.empty()) return true;
        std::ofstream of(path, std::ios::binary);
        if (!of) return false;
        auto u32 = [&](uint32_t u) { of.write((const char*)&u, 4); };
        of.write("TRKSS", 5);
        of.put('a');
        u32(ntracks);
        u32(1);
        u32((uint32_t)description.size());
        of.write(description.data(), (std::streamsize)description.size());
        u32(times.front());
        u32(times.back());
        u32((uint32_t)times.size());
        for (size_t f = 0; f < times.size(); ++f) {
            u32(times[f]);
            of.write((const char*)&xyz[f * 3 * ntracks], (std::streamsize)(3 * ntracks * sizeof(float)));
        }
        times.clear();
        xyz.clear();
        return (bool)of;
    }
};

// ---- output/datapoint_writer.cu: "t<TAB>val" lines after a three-line header, buffered ----
struct PointWriter {
    std::string path;
    std::vector<float> t, val;
    size_t cap = 128;  // FILENAME_LEN entries (main.cu:531)
    PointWriter(const std::string& p, const std::string& title) : path(p) {
        std::ofstream of(path);
        of << "# Data point writer\n# " << title << "\n# Data ordering: t, val" << std::endl;
    }
    void flush() {
        if (t.empty()) return;
        std::ofstream of(path, std::ios::app);
        of.precision(6);
        for (size_t i = 0; i < t.size(); ++i) of << t[i] << '\t' << val[i] << std::endl;
        t.clear();
        val.clear();
    }
    void append(float time, float v) {
        t.push_back(time);
        val.push_back(v);
        if (t.size() >= cap) flush();
    }
    ~PointWriter() { flush(); }
};

std::string fmt(const char* f, double a) {
    char b[256];
    snprintf(b, sizeof b, f, a);
    return b;
}
std::string fmti(const char* f, int a) {
    char b[256];
    snprintf(b, sizeof b, f, a);
    return b;
}

int die(const std::string& msg) {
    std::cerr << "Error: " << msg << std::endl;
    return 1;
}

struct Ctx {
    rbc_ctx* c = nullptr;
    std::vector<rbc_tracer*> tr;
    ~Ctx() {
        for (rbc_tracer* t : tr) rbc_tracer_destroy(t);
        if (c) rbc_destroy(c);
    }
};

}  // namespace

extern "C" int rbc_run_main(int argc, const char* const* argv) {
    Args A;
    for (int i = 0; i < argc; ++i) A.v.push_back(argv[i] ? argv[i] : "");
    auto bad = [&]() { return die(A.err); };

    // precision first: it decides the real type the other arguments are read in
    std::string prec = "f32";
    {
        std::vector<std::string> s;
        if (A.strings(s, 1, "-precision")) prec = s[0];
    }
    int precision;
    if (prec == "f32") precision = RBC_F32;
    else if (prec == "f64") precision = RBC_F64;
    else if (prec == "bf16x3") precision = RBC_BF16X3;
    else return die("-precision must be f32, f64 or bf16x3");
    const bool fp32 = precision != RBC_F64;
    // a real argument in the build's real type (CUDA_FLOAT_REAL), returned as double
    auto real = [&](double& out, const char* name, bool mandatory) {
        if (fp32) {
            float f = (float)out;
            if (!A.get1<float>(f, name, mandatory)) return false;
            out = f;
            return true;
        }
        return A.get1<double>(out, name, mandatory);
    };
    auto reals = [&](std::vector<double>& out, int count, const char* name, bool mandatory) {
        if (fp32) {
            std::vector<float> f;
            if (!A.get<float>(f, count, name, mandatory)) return false;
            out.assign(f.begin(), f.end());
            return true;
        }
        return A.get<double>(out, count, name, mandatory);
    };

    int device = 0;
    if (!A.get1<int>(device, "-device", false)) return bad();
    std::string dir;
    {
        std::vector<std::string> s;
        if (A.strings(s, 1, "-dir")) dir = s[0];
    }
    std::vector<int> dims;
    if (!A.get<int>(dims, 3, "-nmodes", true)) return bad();
    const int M = dims[0], N = dims[1], L = dims[2];
    std::vector<double> clen;
    if (!reals(clen, 2, "-clength", true)) return bad();

    // the theta start (main.cu:90-96: a later option in this list overrides an earlier one)
    enum { INIT_RANDOM, INIT_TEST, INIT_PATTERN } init = INIT_RANDOM;
    int pattern = 0;
    {
        static const struct { const char* opt; int kind; } pats[] = {
            {"-init_rectangle", RBC_PATTERN_RECTANGLE}, {"-init_dislocation", RBC_PATTERN_DISLOCATION},
            {"-init_isr", RBC_PATTERN_ISR},             {"-init_sv", RBC_PATTERN_SV},
            {"-init_cr", RBC_PATTERN_CR},               {"-init_giantspiral", RBC_PATTERN_GIANTSPIRAL}};
        for (const auto& pt : pats)
            if (A.opt(pt.opt)) {
                init = INIT_PATTERN;
                pattern = pt.kind;
            }
    }
    if (A.opt("-init_test")) init = INIT_TEST;
    double init_mode = clen[0] / 2.0;
    if (!A.get1<double>(init_mode, "-init_mode", false)) return bad();

    double Ra = 1900, Pr = 1.0, dt = 0.001, final_t = 0.001;
    if (!real(Ra, "-ra", true) || !real(Pr, "-pr", true)) return bad();
    // -dt / -t (main.cu:108-109); README.txt:38-39 names them -delta_t / -final_t
    if (!real(dt, A.opt("-dt") || !A.opt("-delta_t") ? "-dt" : "-delta_t", true)) return bad();
    if (!real(final_t, A.opt("-t") || !A.opt("-final_t") ? "-t" : "-final_t", true)) return bad();
    int nfiles = 1;
    if (!A.get1<int>(nfiles, "-nfiles", false)) return bad();
    double eta = 0.01;
    if (!real(eta, "-eta", false)) return bad();
    // -penalization <shape> runs what the PENALIZED builds compile in (run/Makefile: the default
    // rectangle, -DCIRCLE, -DDISKARRAY) with the -eta given; without it eta is unused, as in the
    // reference's default build (main.cu:120-122)
    int pen_shape = -1;
    {
        std::vector<std::string> sh;
        if (A.strings(sh, 1, "-penalization")) {
            if (sh[0] == "rectangle") pen_shape = RBC_MASK_RECTANGLE;
            else if (sh[0] == "circle") pen_shape = RBC_MASK_CIRCLE;
            else if (sh[0] == "diskarray") pen_shape = RBC_MASK_DISKARRAY;
            else if (sh[0] == "everywhere") pen_shape = RBC_MASK_EVERYWHERE;
            else return die("-penalization must be rectangle, circle, diskarray or everywhere");
        }
    }
    if (pen_shape < 0) eta = 0;
    std::vector<std::string> infiles;
    const bool have_files = A.strings(infiles, 5, "-files");
    std::vector<std::string> histfiles;
    const bool have_hist = A.strings(histfiles, 5, "-history_files");
    int seed = 0;
    if (!A.get1<int>(seed, "-seed", false)) return bad();
    double mean_F = 0.0;
    if (!real(mean_F, "-mean_F", false)) return bad();

    // particles (main.cu:141-197)
    int npart = 0;
    if (!A.get1<int>(npart, "-npart", false)) return bad();
    int ptype = RBC_PART_RANDOM;
    double psig = 0.0, pcx = 0.0, pcy = 0.0;
    if (A.opt("-part_center")) ptype = RBC_PART_CENTER;
    if (A.opt("-part_center_sigma")) {
        ptype = RBC_PART_CENTER_SIGMA;
        psig = 0.125;
    }
    if (A.opt("-part_diagonal")) ptype = RBC_PART_DIAGONAL;
    if (A.opt("-part_pairs")) ptype = RBC_PART_PAIRS;
    if (A.opt("-part_coord_sigma")) {
        ptype = RBC_PART_COORD_SIGMA;
        std::vector<double> co;
        if (!reals(co, 2, "-part_coord_sigma", true)) return bad();
        pcx = co[0];
        pcy = co[1];
        psig = 0.125;
    }
    if (A.opt("-part_sigma") && !real(psig, "-part_sigma", true)) return bad();
    const bool write_tracks = A.opt("-write_tracks_every");
    int track_period = 1;
    if (!A.get1<int>(track_period, "-write_tracks_every", false)) return bad();
    const bool calc_m2 = A.opt("-calc_secondmoment");
    int le_count = 1;
    if (!A.get1<int>(le_count, "-le_count", false)) return bad();
    std::vector<double> lewis;
    if (A.opt("-le") && !reals(lewis, le_count, "-le", false)) return bad();
    std::string tmode = "rk4";
    {
        std::vector<std::string> s;
        if (A.strings(s, 1, "-tracer_mode")) tmode = s[0];
    }
    if (tmode != "rk4" && tmode != "spectral") return die("-tracer_mode must be rk4 or spectral");

    int snapshot_freq = 0;
    if (!A.get1<int>(snapshot_freq, "-snapshot_freq", false)) return bad();
    const bool do_step = !A.opt("-no_step_time");
    int freeze_circle = 0, temp_circle = 0, port = 0;
    if (!A.get1<int>(freeze_circle, "-freeze_circle", false) || !A.get1<int>(temp_circle, "-temp_circle", false) ||
        !A.get1<int>(port, "-port", false))
        return bad();
    const bool write_hist = A.opt("-write_history");
    int nq = 0;
    if (!A.get1<int>(nq, "-nq", false)) return bad();

    // working directory (main.cu:234-252): "" -> ../data, absolute paths as given, else
    // ../data/<dir>; a path prefix stands in for the chdir
    std::string wd;
    if (dir.empty()) wd = "../data";
    else if (dir[0] == '/' || dir[0] == '~') wd = dir;
    else wd = "../data/" + dir;
    mkdir(wd.c_str(), 0777);
    struct stat sb;
    if (stat(wd.c_str(), &sb) != 0 || !S_ISDIR(sb.st_mode))
        return die("not able to create and change to output directory!");
    auto at = [&](const std::string& name) { return (!name.empty() && name[0] == '/') ? name : wd + "/" + name; };

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess) ndev = 0;
    std::cout << "number of GPUs: " << ndev << std::endl;
    if (device >= ndev || device < 0) return die("please select a valid GPU device!");
    hipDeviceProp_t prop;
    std::string devname = "?";
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) devname = prop.name;
    std::cout << "device name: " << devname << std::endl;
    std::cout << "number of spectral coefficients: " << M / 2 + 1 << " x " << N << " x " << L << std::endl;

    // number of iterations in the build's real type (CUDA_FLOAT_REAL division, main.cu:279)
    const int iterations = fp32 ? iceil((double)((float)final_t / (float)dt)) : iceil(final_t / dt);
    const int write_file = nfiles > 0 ? iterations / nfiles : 0;
    std::cout << "number of iterations: " << iterations << " with dt " << dt << std::endl;
    std::cout << "writing file each " << write_file << "th iteration" << std::endl;

    Ctx X;
    rbc_params p{};
    p.M = M;
    p.N = N;
    p.L = L;
    p.Lx = clen[0];
    p.Ly = clen[1];
    p.Ra = Ra;
    p.Pr = Pr;
    p.dt = dt;
    p.precision = precision;
    p.nq = nq;
    p.device = device;
    if (rbc_create(&p, &X.c) != RBC_OK) return die(rbc_last_error(nullptr));
    rbc_ctx* c = X.c;
    auto cfail = [&](const char* what) { return die(std::string(what) + ": " + rbc_last_error(c)); };

    auto paths5 = [&](const std::vector<std::string>& names, std::vector<std::string>& full,
                      std::vector<const char*>& ptrs) {
        full.clear();
        for (const auto& n : names) full.push_back(at(n));
        ptrs.clear();
        for (const auto& f : full) ptrs.push_back(f.c_str());
    };
    if (have_files) {
        std::cout << "...reading init data from binary file" << std::endl;
        std::vector<std::string> full;
        std::vector<const char*> ptr;
        paths5(infiles, full, ptr);
        if (rbc_read_state_files(c, ptr.data(), 0) != RBC_OK) return cfail("reading the state files");
        std::cout << "...reading init data from binary file finished" << std::endl;
    } else if (init == INIT_PATTERN) {
        if (rbc_init_pattern(c, pattern, init_mode, (unsigned long long)seed) != RBC_OK) return cfail("initial pattern");
    } else if (rbc_init_state(c, init == INIT_TEST ? RBC_INIT_TEST : RBC_INIT_RANDOM, (unsigned long long)seed,
                              1e-6) != RBC_OK) {
        return cfail("initial state");
    }
    if (have_hist) {
        std::vector<std::string> full;
        std::vector<const char*> ptr;
        paths5(histfiles, full, ptr);
        if (rbc_read_state_files(c, ptr.data(), 1) != RBC_OK) return cfail("reading the history files");
    }
    if (pen_shape >= 0 && rbc_set_penalization(c, eta, pen_shape, nullptr) != RBC_OK) return cfail("penalization");

    {  // simulation.log (main.cu:401-424)
        std::ofstream log(at("simulation.log"));
        for (int k = 0; k < 2; ++k) {
            std::ostream& s = k == 0 ? (std::ostream&)log : std::cout;
            s << "GPU device " << device << ": " << devname << "\n";
            s << "Simulation of Rayleigh-Benard convection\n";
            s << "Rayleigh-Number: " << Ra << "\n";
            s << "Prandtl-Number: " << Pr << "\n";
            s << "delta t: " << dt << "\n";
            s << "final t: " << final_t << "\n";
            s << "number of files: " << nfiles << "\n";
            s << "number of iterations: " << iterations << "\n";
            s << "eta(penalization): " << eta << "\n";
            s << "cube length x direction: " << clen[0] << "\n";
            s << "cube length y direction: " << clen[1] << "\n";
            s << "cube length z direction: " << 1.0 << "\n";
            s << "number of x modes: " << M << "\n";
            s << "number of y modes: " << N << "\n";
            s << "number of z modes: " << L << "\n";
            s << "precision: " << prec << "\n";
            s << "command string: ";
            for (int i = 0; i < argc; ++i) s << A.v[i] << ' ';
            s << std::endl;
        }
    }

    // tracers, one per Lewis number (main.cu:491-520); they exist only with -le (main.cu:494)
    const int kRing = 256;
    std::vector<TrackWriter> tracks;
    std::vector<std::unique_ptr<PointWriter>> m2w;
    if (npart > 0) {
        for (size_t k = 0; k < lewis.size(); ++k) {
            rbc_tracer* t = nullptr;
            if (rbc_tracer_create(c, npart, kRing, tmode == "rk4" ? RBC_TRACER_RK4_TRILINEAR : RBC_TRACER_SPECTRAL_RK2,
                                  lewis[k], (unsigned long long)k, &t) != RBC_OK)
                return cfail("tracer");
            X.tr.push_back(t);
            if (rbc_tracer_init(t, ptype, pcx, pcy, 0.0, psig) != RBC_OK) return cfail("tracer positions");
            if (write_tracks) {
                char d[512];
                snprintf(d, sizeof d,
                         "Particle tracks.  Simulation: Ra=%.1lf Pr=%.1lf Size=%.1lfx%.1lfx1.0  Technical: "
                         "Modes=%dx%dx%d dt=%.3lf  Tracks: Le=%.0le ",
                         Ra, Pr, clen[0], clen[1], M, N, L, dt, lewis[k]);
                tracks.push_back(TrackWriter());
                tracks.back().description = d;
                tracks.back().period = track_period;
            }
            if (calc_m2)
                m2w.emplace_back(new PointWriter(at(fmti("secondmoment_%d.dat", (int)k)), fmt("Le %.9f", lewis[k])));
        }
    }
    const int ntr = (int)X.tr.size();
    std::vector<float> bx, by, bz;
    if (ntr) {
        bx.resize((size_t)(kRing - 1) * npart);
        by.resize(bx.size());
        bz.resize(bx.size());
    }

    std::cout << "Start time loop" << std::endl;
    const bool per_step = ntr > 0 || mean_F != 0.0 || freeze_circle || temp_circle;
    for (int i = 1; i <= iterations;) {
        // steps until the next output (one at a time when anything happens between steps)
        int n = 1;
        if (!per_step && do_step) {
            const int next = write_file > 0 ? ((i + write_file - 1) / write_file) * write_file : iterations;
            n = std::max(1, std::min(next, iterations) - i + 1);
        }
        if (do_step) {
            if (mean_F != 0.0 && rbc_force_mean(c, mean_F) != RBC_OK) return cfail("force_mean");
            if (freeze_circle) {  // main.cu:569-577: the operator, then the capture at step n
                if (rbc_apply_freeze(c) != RBC_OK) return cfail("freeze operator");
                if (i == freeze_circle && rbc_freeze(c, RBC_PLANE_CIRCLE, 0.8) != RBC_OK) return cfail("freeze_circle");
            }
            if (temp_circle) {  // main.cu:580-588: the heat source, prepared (a rectangle) at step n
                if (rbc_apply_temp_mask(c) != RBC_OK) return cfail("heat source");
                if (i == temp_circle && rbc_temp_mask(c, RBC_PLANE_RECTANGLE, 0.1) != RBC_OK)
                    return cfail("temp_rectangle");
            }
            if (rbc_step(c, n) != RBC_OK) return cfail("step");
        }
        const int last = i + n - 1;
        for (int k = 0; k < ntr; ++k) {
            rbc_tracer* t = X.tr[k];
            if (rbc_tracer_step(t, dt) != RBC_OK) return cfail("tracer step");
            int it = 0, ptime = 0;
            rbc_tracer_info(t, &it, &ptime);
            if (it >= kRing - 1) {  // is_buffer_full
                std::cout << "read out particle positions" << std::endl;
                int nt = 0;
                if (rbc_tracer_get_positions(t, 1, bx.data(), by.data(), bz.data(), &nt) != RBC_OK)
                    return cfail("particle positions");
                if (write_tracks) tracks[k].append(bx.data(), by.data(), bz.data(), npart, nt);
            }
            if (calc_m2) {
                double mom[4];
                if (rbc_tracer_moments(t, mom) != RBC_OK) return cfail("second moment");
                m2w[k]->append((float)(last * dt), (float)mom[2]);
            }
        }
        if (write_file > 0 && last % write_file == 0) {
            if (rbc_synchronize(c) != RBC_OK) return cfail("time step");
            std::cout << "writing tmp-files to disk!" << std::endl;
            const char* names[5] = {"theta", "f", "g", "F", "G"};
            std::vector<std::string> full;
            std::vector<const char*> ptr;
            for (const char* nm : names) full.push_back(at(std::string(nm) + fmti("_%08d.bin", last)));
            for (const auto& f : full) ptr.push_back(f.c_str());
            if (rbc_write_state_files(c, ptr.data(), 0, 0) != RBC_OK) return cfail("writing the state files");
            if (write_hist && do_step) {
                full.clear();
                ptr.clear();
                for (const char* nm : names) full.push_back(at(std::string("hist_") + nm + fmti("_%08d.bin", last)));
                for (const auto& f : full) ptr.push_back(f.c_str());
                if (rbc_write_state_files(c, ptr.data(), 1, 0) != RBC_OK) return cfail("writing the history files");
            }
            std::cout << "wrote file" << std::endl;
            for (int k = 0; k < (int)tracks.size(); ++k) {
                char nm[128];
                snprintf(nm, sizeof nm, "particles_le%.0le_%08d.trkss", lewis[k], last);
                if (!tracks[k].write(at(nm))) return die(std::string("Open file failed: ") + nm);
            }
        }
        i = last + 1;
    }
    if (rbc_synchronize(c) != RBC_OK) return cfail("time step");
    return 0;
}

// ≙ main() of gpu_flow/code/main_extract_vel.cu: reads a state (-files, five paths; the grid
// from the theta file's header) and, with -write_charvel_nz > 0, writes the characteristic
// velocity <|u|> over 10 heights (the reference tool's fixed count, :190) to -outfile_charvel
// as a data-point file ("0<TAB>U").  The gnuplot velocity and temperature slices
// (-write_vel_nz, -write_temp_nz) are not available in this build.
extern "C" int rbc_extract_vel_main(int argc, const char* const* argv) {
    Args A;
    for (int i = 0; i < argc; ++i) A.v.push_back(argv[i] ? argv[i] : "");
    int device = 0;
    if (!A.get1<int>(device, "-device", false)) return die(A.err);
    std::vector<float> clen;
    if (!A.get<float>(clen, 2, "-clength", true)) return die(A.err);
    std::vector<std::string> files;
    if (!A.strings(files, 5, "-files")) return die("Mandatory argument not found: -files");
    int vel_nz = 0, temp_nz = 0, charvel_nz = 0;
    if (!A.get1<int>(vel_nz, "-write_vel_nz", false) || !A.get1<int>(temp_nz, "-write_temp_nz", false) ||
        !A.get1<int>(charvel_nz, "-write_charvel_nz", false))
        return die(A.err);
    if (vel_nz > 0) return die("-write_vel_nz is not available in this build");
    if (temp_nz > 0) return die("-write_temp_nz is not available in this build");
    std::vector<std::string> out;
    if (charvel_nz > 0 && !A.strings(out, 1, "-outfile_charvel"))
        return die("Mandatory argument not found: -outfile_charvel");
    std::vector<std::string> descr;
    A.strings(descr, 1, "-description");

    // dims from the theta file: int32 count, int32 ndims, ndims int32 (mox, N, L)
    int hdr[5] = {0, 0, 0, 0, 0};
    {
        std::ifstream in(files[0], std::ios::binary);
        if (!in.read((char*)hdr, sizeof hdr) || hdr[0] != 1 || hdr[1] != 3)
            return die("not a state file: " + files[0]);
    }
    rbc_params p{};
    p.M = 2 * (hdr[2] - 1);
    p.N = hdr[3];
    p.L = hdr[4];
    p.Lx = clen[0];
    p.Ly = clen[1];
    p.Ra = 1.0;
    p.Pr = 1.0;
    p.dt = 1.0;
    p.precision = RBC_F32;
    p.device = device;
    std::cout << "number of spectral coefficients: " << p.M / 2 + 1 << " x " << p.N << " x " << p.L << std::endl;
    Ctx X;
    if (rbc_create(&p, &X.c) != RBC_OK) return die(rbc_last_error(nullptr));
    std::vector<const char*> ptr;
    for (const auto& f : files) ptr.push_back(f.c_str());
    if (rbc_read_state_files(X.c, ptr.data(), 0) != RBC_OK) return die(rbc_last_error(X.c));
    if (charvel_nz > 0) {
        const int nz = 10;
        std::cout << "Extract characteristic velocity U on " << charvel_nz << " levels." << std::endl;
        double U = 0.0;
        if (rbc_characteristic_velocity(X.c, nz, &U) != RBC_OK) return die(rbc_last_error(X.c));
        std::cout << "U=" << U << std::endl;
        PointWriter w(out[0], descr.empty() ? std::string() : descr[0]);
        w.append(0.0f, (float)U);
    }
    std::cout << "finished write\n" << std::endl;
    return 0;
}
