// The reference driver's .conf front end on the engine (see include/hawaii_conf.h).
//
// Parsing follows the reference's grammar (src/driver.c:54-1522): blocks opened by &INPUT,
// &MONOMER or &PROCESSING and closed by &END; `KEYWORD = value` with case-insensitive
// keywords, values being numbers, bare words, double-quoted strings or {a, b, c} lists;
// `!` starts a comment; PROCESSING holds calls NAME(args).
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <cctype>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/hawaii_conf.h"
#include "hh_conf_impl.h"
#include "hh_internal.h"

namespace hh {
namespace conf {

using hh::fail;

constexpr double kALU = 5.29177210903e-11;                    // src/constants.h:8
constexpr double kZeroCoeff = 0.00361479637 / (4.0 * M_PI);   // src/constants.h:29
constexpr double kSecondCoeff = 13856114.29344114;              // src/constants.h:30

struct Tok {
    enum Kind { END, IDENT, NUMBER, STRING, SYM, BLOCK } kind = END;
    std::string text;
    double num = 0.0;
    int line = 0;
};

class Lexer {
  public:
    explicit Lexer(const std::string& s) : s_(s) {}
    Tok next()
    {
        skip();
        Tok t;
        t.line = line_;
        if (i_ >= s_.size()) return t;
        const char c = s_[i_];
        if (c == '&') {
            size_t j = ++i_;
            while (j < s_.size() && (std::isalnum((unsigned char)s_[j]) || s_[j] == '_')) ++j;
            t.kind = Tok::BLOCK;
            t.text = upper(s_.substr(i_, j - i_));
            i_ = j;
            return t;
        }
        if (c == '"') {
            size_t j = i_ + 1;
            while (j < s_.size() && s_[j] != '"' && s_[j] != '\n') ++j;
            t.kind = Tok::STRING;
            t.text = s_.substr(i_ + 1, j - i_ - 1);
            i_ = j < s_.size() && s_[j] == '"' ? j + 1 : j;
            return t;
        }
        if (std::isdigit((unsigned char)c) || c == '-' || c == '+' || c == '.') {
            // numbers may group digits with underscores (100_000_000)
            size_t j = i_;
            std::string clean;
            while (j < s_.size()) {
                const char d = s_[j];
                const bool sign_ok = (j == i_) || s_[j - 1] == 'e' || s_[j - 1] == 'E';
                if (std::isdigit((unsigned char)d) || d == '.' || d == 'e' || d == 'E' || ((d == '+' || d == '-') && sign_ok)) {
                    clean += d;
                } else if (d != '_') {
                    break;
                }
                ++j;
            }
            char* e = nullptr;
            const double v = std::strtod(clean.c_str(), &e);
            if (!clean.empty() && e == clean.c_str() + clean.size()) {
                t.kind = Tok::NUMBER;
                t.num = v;
                t.text = s_.substr(i_, j - i_);
                i_ = j;
                return t;
            }
        }
        if (std::isalpha((unsigned char)c) || c == '_') {
            size_t j = i_;
            while (j < s_.size() && (std::isalnum((unsigned char)s_[j]) || s_[j] == '_' || s_[j] == '.')) ++j;
            t.kind = Tok::IDENT;
            t.text = s_.substr(i_, j - i_);
            i_ = j;
            return t;
        }
        t.kind = Tok::SYM;
        t.text = std::string(1, c);
        ++i_;
        return t;
    }
    static std::string upper(std::string x)
    {
        for (char& ch : x) ch = (char)std::toupper((unsigned char)ch);
        return x;
    }

  private:
    void skip()
    {
        while (i_ < s_.size()) {
            const char c = s_[i_];
            if (c == '\n') {
                ++line_;
                ++i_;
            } else if (std::isspace((unsigned char)c)) {
                ++i_;
            } else if (c == '!') {
                while (i_ < s_.size() && s_[i_] != '\n') ++i_;
            } else {
                break;
            }
        }
    }
    const std::string& s_;
    size_t i_ = 0;
    int line_ = 1;
};

// keywords of the reference that steer parts outside this engine: accepted, ignored
const char* kIgnored[] = {"ODD_J_SPIN_WEIGHT", "EVEN_J_SPIN_WEIGHT", "DJ", "INITIAL_J", "TORQUE_CACHE_LEN", "TORQUE_LIMIT",
                          "NSWITCH_HISTOGRAM_BINS", "NSWITCH_HISTOGRAM_MAX", "NSWITCH_HISTOGRAM_FILENAME",
                          "JINI_HISTOGRAM_BINS", "JINI_HISTOGRAM_MAX", "JINI_HISTOGRAM_FILENAME",
                          "JFIN_HISTOGRAM_BINS", "JFIN_HISTOGRAM_MAX", "JFIN_HISTOGRAM_FILENAME"};

bool ignored(const std::string& k)
{
    for (const char* x : kIgnored)
        if (k == x) return true;
    return false;
}

void copy_str(char* dst, const std::string& s) { snprintf(dst, HH_CONF_PATH, "%s", s.c_str()); }

int parse_text(const std::string& text, const char* path, hh_conf* c)
{
    std::memset(c, 0, sizeof *c);
    c->niterations = 1;
    Lexer lx(text);
    std::string block;
    Tok t = lx.next();
    auto err = [&](int line, const char* fmt, const std::string& what) -> int {
        char msg[256];
        snprintf(msg, sizeof msg, fmt, what.c_str());
        return fail(HH_E_INVALID, "%s:%d: %s", path, line, msg);
    };
    while (t.kind != Tok::END) {
        if (t.kind == Tok::BLOCK) {
            if (t.text == "END") {
                if (block.empty()) return err(t.line, "%s without an open block", "&END");
                block.clear();
            } else if (t.text == "INPUT" || t.text == "PROCESSING") {
                block = t.text;
            } else if (t.text == "MONOMER") {
                if (c->n_monomers >= 2) return err(t.line, "%s: at most two monomers", "&MONOMER");
                block = t.text;
                ++c->n_monomers;
            } else {
                return err(t.line, "unknown block &%s", t.text);
            }
            t = lx.next();
            continue;
        }
        if (block.empty()) return err(t.line, "'%s' outside of a block", t.text);
        if (t.kind != Tok::IDENT) return err(t.line, "unexpected '%s'", t.text);
        const std::string key = Lexer::upper(t.text);
        const int line = t.line;

        if (block == "PROCESSING") {
            // NAME(arg, ...), or a setting KEY = value of the processing stack
            Tok open = lx.next();
            if (open.text == "=") {
                const Tok val = lx.next();
                if (key == "SPECTRUM_FREQUENCY_MAX") {
                    if (val.kind != Tok::NUMBER) return err(line, "%s expects a number", key);
                    c->spectrum_frequency_max = val.num;
                } else if (!ignored(key)) {
                    return err(line, "unknown PROCESSING setting %s", key);
                }
                t = lx.next();
                continue;
            }
            if (open.text != "(") return err(line, "expected '(' after %s", key);
            std::vector<Tok> args;
            Tok a = lx.next();
            while (a.kind != Tok::END && a.text != ")") {
                if (a.text != ",") args.push_back(a);
                a = lx.next();
            }
            if (c->nproc >= HH_CONF_MAX_PROC) return err(line, "too many PROCESSING calls (%s)", key);
            hh_conf_proc& pr = c->proc[c->nproc];
            static const struct { const char* name; int op; int nargs; } ops[] = {
                {"READ_CF", HH_PROC_READ_CF, 1}, {"READ_SF", HH_PROC_READ_SF, 1}, {"CF_TO_SF", HH_PROC_CF_TO_SF, 0},
                {"CMP", HH_PROC_CMP, 0},          {"DUP", HH_PROC_DUP, 0},         {"DUP2", HH_PROC_DUP2, 0},
                {"DROP", HH_PROC_DROP, 0},        {"DROP2", HH_PROC_DROP2, 0},     {"WRITE_CF", HH_PROC_WRITE_CF, 1},
                {"WRITE_SF", HH_PROC_WRITE_SF, 1}, {"READ_SPECTRUM", HH_PROC_READ_SPECTRUM, 1},
                {"WRITE_SPECTRUM", HH_PROC_WRITE_SPECTRUM, 1}, {"ALPHA", HH_PROC_ALPHA, 0}, {"D1", HH_PROC_D1, 0},
                {"D2", HH_PROC_D2, 0}, {"D3", HH_PROC_D3, 0}, {"D4", HH_PROC_D4, 0}, {"D4A", HH_PROC_D4A, 0},
                {"AVERAGE_CFS", HH_PROC_AVERAGE_CFS, 0}, {"INT3", HH_PROC_INT3, 0}};
            if (key == "COMPUTE_MN_CLASSICAL_DETAILED_BALANCE" || key == "COMPUTE_MN_QUANTUM_DETAILED_BALANCE") {
                // one integer argument, positional or named n (src/driver.c:2304-2330)
                const bool named = args.size() == 3 && Lexer::upper(args[0].text) == "N" && args[1].text == "=";
                if (!(named || args.size() == 1)) return err(line, "%s expects one argument n", key);
                const Tok& v = args[named ? 2 : 0];
                if (v.kind != Tok::NUMBER || v.num != (double)(long)v.num)
                    return err(line, "%s expects an integer order n", key);
                snprintf(pr.arg, sizeof pr.arg, "%ld", (long)v.num);
                pr.op = key == "COMPUTE_MN_QUANTUM_DETAILED_BALANCE" ? HH_PROC_MN_QUANTUM : HH_PROC_MN_CLASSICAL;
                ++c->nproc;
                t = lx.next();
                continue;
            }
            if (key == "FIT_BASELINE") {
                // at most one integer argument, positional or named cf_extrapolation_begin_index
                // (src/driver.c:2096-2118); kept as its value, -1 for the default
                long begin = -1;
                if (!args.empty()) {
                    const bool named = args.size() == 3 && args[1].text == "=";
                    if (!(named || args.size() == 1))
                        return err(line, "function call %s expects at most 1 arguments", key);
                    if (named && Lexer::upper(args[0].text) != "CF_EXTRAPOLATION_BEGIN_INDEX")
                        return err(line, "function call FIT_BASELINE expects a named argument "
                                         "cf_extrapolation_begin_index but got %s", args[0].text);
                    const Tok& v = args[named ? 2 : 0];
                    if (v.kind != Tok::NUMBER || v.num != (double)(long)v.num || v.num < 0)
                        return err(line, "%s expects a non-negative integer index", key);
                    begin = (long)v.num;
                }
                snprintf(pr.arg, sizeof pr.arg, "%ld", begin);
                pr.op = HH_PROC_FIT_BASELINE;
                ++c->nproc;
                t = lx.next();
                continue;
            }
            if (key == "SMOOTH") {
                // named arguments only (src/driver.c:2879-2909), kept as
                // "grid_npoints;grid_max;window_size_min;window_size_step;window_size_delay"
                long npoints = 0, ws_min = 21, ws_delay = 10;
                double grid_max = 0.0, ws_step = 0.1;
                for (size_t k = 0; k < args.size(); k += 3) {
                    if (k + 2 >= args.size() || args[k + 1].text != "=" || args[k + 2].kind != Tok::NUMBER)
                        return err(line, "function call %s accepts only named arguments", key);
                    const std::string name = Lexer::upper(args[k].text);
                    const double v = args[k + 2].num;
                    const bool integer = v == (double)(long)v;
                    if (name == "GRID_NPOINTS" || name == "WINDOW_SIZE_MIN" || name == "WINDOW_SIZE_DELAY") {
                        if (!integer || v < 0) return err(line, "SMOOTH: %s expects a non-negative integer", args[k].text);
                        (name == "GRID_NPOINTS" ? npoints : name == "WINDOW_SIZE_MIN" ? ws_min : ws_delay) = (long)v;
                    } else if (name == "GRID_MAX") {
                        grid_max = v;
                    } else if (name == "WINDOW_SIZE_STEP") {
                        ws_step = v;
                    } else {
                        // reported and skipped, as the reference does
                        std::fprintf(stderr, "ERROR: line %d: function call %s got unexpected named argument %s\n", line,
                                     key.c_str(), args[k].text.c_str());
                    }
                }
                snprintf(pr.arg, sizeof pr.arg, "%ld;%.17g;%ld;%.17g;%ld", npoints, grid_max, ws_min, ws_step, ws_delay);
                pr.op = HH_PROC_SMOOTH;
                ++c->nproc;
                t = lx.next();
                continue;
            }
            if (key == "ADD_SPECTRA") {
                // named arguments only (n = <int>, wavenumber_step = <float>), kept as "n;step"
                long n = -1;
                double step = 0.0;
                for (size_t k = 0; k < args.size(); k += 3) {
                    if (k + 2 >= args.size() || args[k + 1].text != "=" || args[k + 2].kind != Tok::NUMBER)
                        return err(line, "function call %s accepts only named arguments", key);
                    const std::string name = Lexer::upper(args[k].text);
                    if (name == "N")
                        n = (long)args[k + 2].num;
                    else if (name == "WAVENUMBER_STEP")
                        step = args[k + 2].num;
                    else
                        return err(line, "function call ADD_SPECTRA got unexpected named argument %s", args[k].text);
                }
                snprintf(pr.arg, sizeof pr.arg, "%ld;%.17g", n, step);
                pr.op = HH_PROC_ADD_SPECTRA;
                ++c->nproc;
                t = lx.next();
                continue;
            }
            int found = -1;
            for (const auto& o : ops)
                if (key == o.name) {
                    found = o.op;
                    if ((int)args.size() != o.nargs || (o.nargs == 1 && args[0].kind != Tok::STRING))
                        return err(line, "%s: wrong arguments", key);
                    if (o.nargs == 1) copy_str(pr.arg, args[0].text);
                }
            if (found < 0) return err(line, "PROCESSING function %s is not part of this engine", key);
            pr.op = found;
            ++c->nproc;
            t = lx.next();
            continue;
        }

        Tok eq = lx.next();
        if (eq.text != "=") return err(line, "expected '=' after %s", key);
        Tok v = lx.next();
        std::vector<double> list;
        std::vector<std::string> slist;
        if (v.text == "{") {
            Tok a = lx.next();
            while (a.kind != Tok::END && a.text != "}") {
                if (a.kind == Tok::NUMBER) list.push_back(a.num);
                if (a.kind == Tok::STRING) slist.push_back(a.text);
                a = lx.next();
            }
        }
        auto num = [&](double& dst) -> int {
            if (v.kind != Tok::NUMBER) return err(line, "%s expects a number", key);
            dst = v.num;
            return 0;
        };
        auto inum = [&](long& dst) -> int {
            if (v.kind != Tok::NUMBER || v.num != std::floor(v.num)) return err(line, "%s expects an integer", key);
            dst = (long)v.num;
            return 0;
        };
        auto str = [&](char* dst) -> int {
            if (v.kind != Tok::STRING) return err(line, "%s expects a double-quoted string", key);
            copy_str(dst, v.text);
            return 0;
        };
        int rc = 0;
        if (block == "MONOMER") {
            const int m = c->n_monomers - 1;
            if (key == "MONOMER_TYPE") {
                const std::string w = Lexer::upper(v.text);
                c->monomer_type[m] = w == "ATOM"              ? HH_MONOMER_ATOM
                                     : w == "LINEAR_MOLECULE" ? HH_MONOMER_LINEAR_MOLECULE
                                     : w == "ROTOR"           ? HH_MONOMER_ROTOR
                                                              : HH_MONOMER_OTHER;
            } else if (key == "II") {
                if (list.size() != 3) return err(line, "%s expects {Ixx, Iyy, Izz}", key);
                for (int k = 0; k < 3; ++k) c->II[m][k] = list[k];
            } else if (!ignored(key)) {
                return err(line, "unknown MONOMER keyword %s", key);
            }
        } else if (key == "CALCULATION_TYPE") {
            const std::string w = Lexer::upper(v.text);
            if (w == "CORRELATION_SINGLE")
                c->calculation_type = HH_CALC_CORRELATION_SINGLE;
            else if (w == "PROCESSING")
                c->calculation_type = HH_CALC_PROCESSING;
            else if (w == "CALCULATE_PHASE_SPACE_M0")
                c->calculation_type = HH_CALC_PHASE_SPACE_M0;
            else if (w == "CALCULATE_PHASE_SPACE_M2")
                c->calculation_type = HH_CALC_PHASE_SPACE_M2;
            else if (w == "PR_MU")
                c->calculation_type = HH_CALC_PR_MU;
            else if (w == "CORRELATION_ARRAY")
                c->calculation_type = HH_CALC_CORRELATION_ARRAY;
            else
                return err(line, "CALCULATION_TYPE %s is not part of this engine", w);
        } else if (key == "PAIR_STATE") {
            const std::string w = Lexer::upper(v.text);
            if (w == "FREE_AND_METASTABLE" || w == "FREE")
                c->pair_state = HH_PAIR_FREE;
            else if (w == "BOUND")
                c->pair_state = HH_PAIR_BOUND;
            else if (w == "ALL")
                c->pair_state = HH_PAIR_ALL;
            else
                return err(line, "unknown PAIR_STATE %s", w);
        } else if (key == "PAIR_REDUCED_MASS") {
            rc = num(c->reduced_mass);
        } else if (key == "SO_POTENTIAL") {
            rc = str(c->so_potential);
        } else if (key == "SO_DIPOLE" || key == "SO_DIPOLE.1") {
            rc = str(c->so_dipole);
        } else if (key == "SO_DIPOLE.2") {
            if (v.kind != Tok::STRING || (c->so_dipole[0] && v.text != c->so_dipole))
                return err(line, "%s: cross-correlation of two dipoles is not part of this engine", key);
        } else if (key == "TEMPERATURE") {
            rc = num(c->temperature);
        } else if (key == "NITERATIONS") {
            rc = inum(c->niterations);
        } else if (key == "TOTAL_TRAJECTORIES") {
            rc = inum(c->total_trajectories);
        } else if (key == "CVODE_TOLERANCE") {
            rc = num(c->cvode_tolerance);
        } else if (key == "SAMPLING_TIME") {
            rc = num(c->sampling_time);
        } else if (key == "MAXTRAJECTORYLENGTH") {
            rc = inum(c->max_trajectory_length);
        } else if (key == "SAMPLER_RMIN") {
            rc = num(c->sampler_Rmin);
        } else if (key == "SAMPLER_RMAX") {
            rc = num(c->sampler_Rmax);
        } else if (key == "PESMIN") {
            rc = num(c->pesmin);
        } else if (key == "INITIALM0_NPOINTS") {
            rc = inum(c->initialM0_npoints);
        } else if (key == "CF_FILENAME") {
            rc = str(c->cf_filename);
        } else if (key == "SF_FILENAME") {
            rc = str(c->sf_filename);
        } else if (key == "R0") {
            rc = num(c->R0);
        } else if (key == "RCUT") {
            rc = num(c->Rcut);
        } else if (key == "PARTIAL_PARTITION_FUNCTION_RATIO") {
            rc = num(c->ppf);
        } else if (key == "HEP_PPF_NITERATIONS") {
            rc = inum(c->hep_ppf_niterations);
        } else if (key == "HEP_PPF_NPOINTS") {
            rc = inum(c->hep_ppf_npoints);
        } else if (key == "HEP_M0_NITERATIONS") {
            rc = inum(c->hep_m0_niterations);
        } else if (key == "HEP_M0_NPOINTS") {
            rc = inum(c->hep_m0_npoints);
        } else if (key == "APPROXIMATEFREQUENCYMAX") {
            rc = num(c->approx_frequency_max);
        } else if (key == "AVERAGE_TIME_BETWEEN_COLLISIONS") {
            rc = num(c->average_time_between_collisions);
        } else if (key == "INITIALM2_NPOINTS") {
            rc = inum(c->initialM2_npoints);
        } else if (key == "HEP_M2_NITERATIONS") {
            rc = inum(c->hep_m2_niterations);
        } else if (key == "HEP_M2_NPOINTS") {
            rc = inum(c->hep_m2_npoints);
        } else if (key == "TEMPERATURES") {
            if (list.empty() || list.size() > 64) return err(line, "%s expects {T1, T2, ...} (at most 64)", key);
            c->ntemperatures = (int)list.size();
            for (size_t k = 0; k < list.size(); ++k) c->temperatures[k] = list[k];
        } else if (key == "SATELLITE_TEMPERATURES") {
            if (list.empty() || list.size() > 64) return err(line, "%s expects {T1, T2, ...} (at most 64)", key);
            c->nsatellite = (int)list.size();
            for (size_t k = 0; k < list.size(); ++k) c->satellite_temperatures[k] = list[k];
        } else if (key == "PARTIAL_PARTITION_FUNCTION_RATIOS") {
            if (list.empty() || list.size() > 64) return err(line, "%s expects {r1, r2, ...} (at most 64)", key);
            c->nppf_ratios = (int)list.size();
            for (size_t k = 0; k < list.size(); ++k) c->ppf_ratios[k] = list[k];
        } else if (key == "CF_FILENAMES") {
            if (slist.empty() || slist.size() > 64) return err(line, "%s expects {\"a\", \"b\", ...} (at most 64)", key);
            c->ncf_filenames = (int)slist.size();
            for (size_t k = 0; k < slist.size(); ++k) copy_str(c->cf_filenames[k], slist[k]);
        } else if (key == "USE_ZIMMERMANN_TRICK") {
            const std::string w = Lexer::upper(v.text);
            c->use_zimmermann_trick = (w == "TRUE" || (v.kind == Tok::NUMBER && v.num != 0));
        } else if (!ignored(key)) {
            return err(line, "unknown keyword %s", key);
        }
        if (rc < 0) return rc;
        t = lx.next();
    }
    if (!block.empty()) return fail(HH_E_INVALID, "%s: block &%s is not closed", path, block.c_str());
    return HH_OK;
}

// ------------------------------------------------------------------ CF / SF files

std::string now_stamp()
{
    time_t raw;
    time(&raw);
    struct tm tmv;
    localtime_r(&raw, &tmv);
    char b[64];
    snprintf(b, sizeof b, "%04d-%02d-%02d %02d:%02d:%02d", tmv.tm_year + 1900, tmv.tm_mon + 1, tmv.tm_mday, tmv.tm_hour,
             tmv.tm_min, tmv.tm_sec);
    return b;
}

// write_correlation_function_ext / write_spectral_function formats (src/hawaii.c:4150-4264)
int write_series(const char* path, const Series& s)
{
    FILE* fp = std::fopen(path, "w");
    if (!fp) return fail(HH_E_INVALID, "could not open '%s' for writing", path);
    std::fprintf(fp, "# HAWAII HYBRID v0.1\n# Saved on %s\n# TEMPERATURE: %.2f\n# AVERAGE OVER %.2f TRAJECTORIES\n",
                 now_stamp().c_str(), s.T, s.ntraj);
    if (s.spectrum) {
        // write_spectrum (src/hawaii.c:4070-4115)
        for (size_t i = 0; i < s.x.size(); ++i) std::fprintf(fp, "%.10f %.10e\n", s.x[i], s.y[i]);
    } else if (s.sf) {
        std::fprintf(fp, "# SPECTRAL FUNCTION LENGTH: %zu\n", s.x.size());
        for (size_t i = 0; i < s.x.size(); ++i) std::fprintf(fp, "%.10f   %.10e\n", s.x[i], s.y[i]);
    } else {
        std::fprintf(fp, "# MAXIMUM TRAJECTORY LENGTH: %zu\n", s.x.size());
        for (size_t i = 0; i < s.x.size(); ++i) std::fprintf(fp, "%.2f %.10e\n", s.x[i], s.y[i]);
    }
    std::fclose(fp);
    return HH_OK;
}

// read_correlation_function / read_spectral_function: TEMPERATURE and AVERAGE OVER from the
// header, then two columns (src/hawaii.c:4856-4939)
int read_series(const char* path, bool sf, Series& s)
{
    std::ifstream in(path);
    if (!in) return fail(HH_E_INVALID, "could not open '%s'", path);
    s = Series{};
    s.sf = sf;
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty()) continue;
        if (line[0] == '#') {
            double v;
            if (std::sscanf(line.c_str(), "# TEMPERATURE: %lf", &v) == 1) s.T = v;
            if (std::sscanf(line.c_str(), "# AVERAGE OVER %lf", &v) == 1) s.ntraj = v;
            continue;
        }
        double a, b;
        if (std::sscanf(line.c_str(), "%lf %lf", &a, &b) == 2) {
            s.x.push_back(a);
            s.y.push_back(b);
        }
    }
    if (s.x.empty()) return fail(HH_E_INVALID, "'%s' holds no data", path);
    return HH_OK;
}

// CMP (src/driver.c execute_cmp): `a` is the element that was on top.  Same length, T within
// 1e-2 K, ntraj within 1e-3, abscissae within 1e-3 a.t.u. (t) / 1e-6 cm^-1 (nu); a value
// differs when |a/b| - 1 > 1e-4 (the reference's one-sided test), or when b is 0 and a is not.
bool cmp_series(const Series& a, const Series& b, std::string& why)
{
    char buf[256];
    if (a.sf != b.sf || a.spectrum != b.spectrum) return why = "incompatible element types", false;
    if (a.x.size() != b.x.size()) {
        snprintf(buf, sizeof buf, "'len' differs (%zu and %zu)", a.x.size(), b.x.size());
        return why = buf, false;
    }
    if (std::fabs(a.T - b.T) > 1e-2) return why = "'Temperature' differs", false;
    if (std::fabs(a.ntraj - b.ntraj) > 1e-3) return why = "'ntraj' differs", false;
    const double xtol = (a.sf || a.spectrum) ? 1e-6 : 1e-3;
    for (size_t i = 0; i < a.x.size(); ++i) {
        if (std::fabs(a.x[i] - b.x[i]) > xtol) {
            snprintf(buf, sizeof buf, "abscissae at index %zu differ (%.6e and %.6e)", i, a.x[i], b.x[i]);
            return why = buf, false;
        }
        const bool differ = b.y[i] != 0.0 ? (std::fabs(a.y[i] / b.y[i]) - 1.0) > 1e-4 : a.y[i] != 0.0;
        if (differ) {
            snprintf(buf, sizeof buf, "values at index %zu differ (%.3e and %.3e)", i, a.y[i], b.y[i]);
            return why = buf, false;
        }
    }
    return true;
}

int cf_to_sf(hh_engine* e, const Series& cf, Series& sf)
{
    const long n = (long)cf.y.size();
    if (n < 2 || (n & (n - 1))) return fail(HH_E_INVALID, "CF_TO_SF: length %ld is not a power of 2", n);
    const double dt = cf.x[1] - cf.x[0];
    double *d = nullptr, *nu = nullptr, *s = nullptr;
    HH_CHECK(hipMalloc(&d, n * sizeof(double)));
    HH_CHECK(hipMalloc(&nu, n * sizeof(double)));
    HH_CHECK(hipMalloc(&s, n * sizeof(double)));
    int rc = HH_OK;
    if (hipMemcpy(d, cf.y.data(), n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(HH_E_DEVICE, "CF_TO_SF: upload failed");
    if (rc == HH_OK) rc = hh_cf_to_sf(e, d, n, dt, nu, s, nullptr);
    sf = Series{};
    sf.sf = true;
    sf.T = cf.T;
    sf.ntraj = cf.ntraj;
    sf.x.resize(n);
    sf.y.resize(n);
    if (rc == HH_OK && (hipMemcpy(sf.x.data(), nu, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
                        hipMemcpy(sf.y.data(), s, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess))
        rc = fail(HH_E_DEVICE, "CF_TO_SF: download failed");
    (void)hipFree(d);
    (void)hipFree(nu);
    (void)hipFree(s);
    return rc;
}

// ------------------------------------------------------------------ desymmetrizations, ALPHA
constexpr double kPlanck = 6.62607015e-34;       // src/constants.h:20
constexpr double kBoltzmann = 1.380649e-23;      // src/constants.h:9
constexpr double kLightSpeedCm = 2.99792458e10;  // src/constants.h:14
constexpr double kHBarSI = 1.054571817e-34;      // src/constants.h:21
constexpr double kATUs = 2.4188843265857e-17;    // src/constants.h:12
constexpr double kADipMomU = 8.4783536255e-30;   // src/constants.h
constexpr double kEpsilon0 = 8.8541878128e-12;   // src/constants.h

double h_nu_over_kT(double nu, double T) { return kPlanck * kLightSpeedCm * nu / (kBoltzmann * T); }

// D1 (desymmetrize_d1, src/hawaii.c:5718-5738), D2 (desymmetrize_d2, src/hawaii.c:5740-5765),
// D3 (desymmetrize_schofield, src/hawaii.c:5696-5716) of an SF
Series desymmetrize(const Series& sf, int which)
{
    Series out = sf;
    for (size_t i = 0; i < sf.x.size(); ++i) {
        const double h = h_nu_over_kT(sf.x[i], sf.T);
        if (which == 1) {
            out.y[i] = sf.y[i] * 2.0 / (1.0 + std::exp(-h));
        } else if (which == 2) {
            out.y[i] = sf.x[i] < 1e-10 ? 0.0 : sf.y[i] * h / (1.0 - std::exp(-h));
        } else {
            out.y[i] = sf.y[i] * std::exp(h / 2.0);
        }
    }
    return out;
}

// Natural cubic spline through (x, y) (the gsl_interp_cspline the reference interpolates with):
// second derivatives m from the tridiagonal system with m_0 = m_{n-1} = 0.
struct NaturalSpline {
    const std::vector<double>& x;
    const std::vector<double>& y;
    std::vector<double> m;
    NaturalSpline(const std::vector<double>& xs, const std::vector<double>& ys) : x(xs), y(ys), m(xs.size(), 0.0)
    {
        const size_t n = x.size();
        if (n < 3) return;
        std::vector<double> diag(n, 0.0), rhs(n, 0.0), up(n, 0.0);
        for (size_t i = 1; i + 1 < n; ++i) {
            const double h0 = x[i] - x[i - 1], h1 = x[i + 1] - x[i];
            diag[i] = 2.0 * (h0 + h1);
            up[i] = h1;
            rhs[i] = 6.0 * ((y[i + 1] - y[i]) / h1 - (y[i] - y[i - 1]) / h0);
        }
        // Thomas algorithm on rows 1 .. n-2 (sub-diagonal h0 of row i = x[i] - x[i-1])
        for (size_t i = 2; i + 1 < n; ++i) {
            const double w = (x[i] - x[i - 1]) / diag[i - 1];
            diag[i] -= w * up[i - 1];
            rhs[i] -= w * rhs[i - 1];
        }
        for (size_t i = n - 2; i >= 1; --i) {
            m[i] = (rhs[i] - (i + 2 < n ? up[i] * m[i + 1] : 0.0)) / diag[i];
            if (i == 1) break;
        }
    }
    double operator()(double t) const
    {
        const size_t n = x.size();
        size_t i = (size_t)(std::upper_bound(x.begin(), x.end(), t) - x.begin());
        i = i == 0 ? 0 : std::min(i - 1, n - 2);
        const double h = x[i + 1] - x[i], A = (x[i + 1] - t) / h, B = 1.0 - A;
        return A * y[i] + B * y[i + 1] + ((A * A * A - A) * m[i] + (B * B * B - B) * m[i + 1]) * h * h / 6.0;
    }
};

// D4 / D4a: Egelstaff's time t -> sqrt(t^2 + (hbar / 2kT)^2) on the CF (egelstaff_time_transform,
// src/hawaii.c:5767-5818; D4a renormalised to the CF's value at t = 0, Frommhold), the kept samples
// zero-padded to a power of two (pad_to_power_of_two, src/hawaii.c:5555-5566), transformed as
// CF_TO_SF and desymmetrized as D3 (desymmetrize_egelstaff_from_cf / _frommhold_from_cf,
// src/hawaii.c:5841-5881).
int egelstaff_sf(hh_engine* e, const Series& cf, bool frommhold, Series& out)
{
    if (!(cf.T > 0)) return fail(HH_E_INVALID, "D4/D4a: the CF needs a temperature");
    const double cc = kHBarSI / kBoltzmann / cf.T / kATUs / 2.0;
    const size_t n = cf.x.size();
    if (n < 3 || !(cc < cf.x[n - 1])) return fail(HH_E_INVALID, "D4/D4a: the CF is shorter than hbar/2kT");
    const NaturalSpline sp(cf.x, cf.y);
    const double norm = frommhold ? sp(0.0) / sp(cc) : 1.0;
    Series eg;
    eg.T = cf.T;
    eg.ntraj = cf.ntraj;
    for (size_t i = 0; i < n; ++i) {
        const double te = std::sqrt(cf.x[i] * cf.x[i] + cc * cc);
        if (te <= cf.x[n - 1]) {
            eg.x.push_back(cf.x[i]);
            eg.y.push_back(norm * sp(te));
        }
    }
    size_t p2 = 1;
    while (p2 < eg.x.size()) p2 <<= 1;
    eg.x.resize(p2, 0.0);
    eg.y.resize(p2, 0.0);
    Series sf;
    const int rc = cf_to_sf(e, eg, sf);
    if (rc < 0) return rc;
    out = desymmetrize(sf, 3);
    return HH_OK;
}

// ALPHA: the absorption coefficient (compute_alpha, src/hawaii.c:5904-5924)
Series alpha(const Series& sf)
{
    const double coeff = (8.0 * M_PI * M_PI * M_PI) * 2.686780111e25 * 2.686780111e25 / 3.0 / kHBarSI;   // Moment_SF_Coeff
    Series sp = sf;
    sp.sf = false;
    sp.spectrum = true;
    for (size_t i = 0; i < sf.x.size(); ++i)
        sp.y[i] = coeff * sp.x[i] * (1.0 - std::exp(-h_nu_over_kT(sp.x[i], sf.T))) * sf.y[i];
    return sp;
}

int run_processing(hh_engine* e, const hh_conf* c, bool quiet)
{
    std::vector<Series> st;
    int code = HH_OK;
    // the wing parameters of the last FIT_BASELINE (Processing_Stack::wing_params, src/driver.c:479):
    // CF_TO_SF adds their Fourier image while any is non-zero, rescaling them in place as the
    // reference does (src/driver.c:1787-1810)
    double wing[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < c->nproc; ++i) {
        const hh_conf_proc& p = c->proc[i];
        auto need = [&](size_t k) -> int {
            if (st.size() < k) return fail(HH_E_INVALID, "PROCESSING call %d needs %zu stack elements", i + 1, k);
            return (int)HH_OK;
        };
        int rc = HH_OK;
        switch (p.op) {
            case HH_PROC_READ_CF:
            case HH_PROC_READ_SF: {
                Series s;
                rc = read_series(p.arg, p.op == HH_PROC_READ_SF, s);
                if (rc == HH_OK) st.push_back(std::move(s));
                break;
            }
            case HH_PROC_CF_TO_SF: {
                if ((rc = need(1)) < 0) break;
                if (st.back().sf) {
                    rc = fail(HH_E_INVALID, "CF_TO_SF expects a correlation function on the stack");
                    break;
                }
                Series sf;
                rc = cf_to_sf(e, st.back(), sf);
                if (rc == HH_OK && (wing[0] != 0.0 || wing[1] != 0.0 || wing[2] != 0.0)) {
                    // the Lorentzian's image, (C + A) / A * A / (2 B) exp(-nu / B), in SF units
                    const double Xs = 1.0 / kLightSpeedCm / kATUs / 2.0 / M_PI;
                    const double Ys = kATUs * kADipMomU * kADipMomU / (4.0 * M_PI * kEpsilon0);
                    wing[0] = wing[0] * Ys * Xs;
                    wing[1] = wing[1] * Xs;
                    wing[2] = wing[2] * Ys * Xs;
                    for (size_t k = 0; k < sf.y.size(); ++k)
                        sf.y[k] += (wing[2] + wing[0]) / wing[0] * (wing[0] / (2.0 * wing[1]) * std::exp(-sf.x[k] / wing[1]));
                }
                if (rc == HH_OK) st.back() = std::move(sf);
                break;
            }
            case HH_PROC_FIT_BASELINE: {
                // execute_fit_baseline (src/driver.c:2072-2142): fit, keep the parameters, subtract
                // the model from the CF on top
                if ((rc = need(1)) < 0) break;
                Series& cf = st.back();
                if (cf.sf || cf.spectrum) {
                    rc = fail(HH_E_INVALID, "FIT_BASELINE expects a correlation function on the stack");
                    break;
                }
                long begin = std::atol(p.arg);
                if (begin < 0) begin = 16384;   // DEFAULT_CF_EXTRAPOLATION_BEGIN_INDEX, src/driver.c:207
                if ((rc = hh_fit_baseline(cf.x.data(), cf.y.data(), (long)cf.y.size(), begin, wing)) < 0) break;
                if (!quiet)
                    std::printf("FIT_BASELINE: A = %.6e, B = %.6e, C = %.6e (from index %ld)\n", wing[0], wing[1], wing[2],
                                begin);
                for (size_t k = 0; k < cf.y.size(); ++k) cf.y[k] -= wing[2] + wing[0] / (1.0 + wing[1] * wing[1] * cf.x[k] * cf.x[k]);
                break;
            }
            case HH_PROC_SMOOTH: {
                // execute_smooth (src/driver.c:2849-2951): the SF on top replaced by its LOESS
                // smoothing on the grid
                if ((rc = need(1)) < 0) break;
                if (!st.back().sf || st.back().spectrum) {
                    rc = fail(HH_E_INVALID, "SMOOTH expects a spectral function on the stack");
                    break;
                }
                long npoints = 0, ws_min = 0, ws_delay = 0;
                double grid_max = 0.0, ws_step = 0.0;
                if (std::sscanf(p.arg, "%ld;%lf;%ld;%lf;%ld", &npoints, &grid_max, &ws_min, &ws_step, &ws_delay) != 5) {
                    rc = fail(HH_E_INVALID, "SMOOTH: bad arguments");
                    break;
                }
                if (npoints == 0) {
                    rc = fail(HH_E_INVALID, "SMOOTH: number of grid points (grid_npoints) for smoothing must be provided");
                    break;
                }
                Series sm;
                sm.sf = true;
                sm.T = st.back().T;
                sm.ntraj = st.back().ntraj;
                sm.x.resize((size_t)npoints);
                sm.y.resize((size_t)npoints);
                rc = hh_loess(st.back().x.data(), st.back().y.data(), (long)st.back().x.size(), grid_max, npoints, ws_min,
                              ws_step, ws_delay, sm.x.data(), sm.y.data());
                if (rc == HH_OK) st.back() = std::move(sm);
                break;
            }
            case HH_PROC_INT3: {
                // execute_int3 (src/driver.c:2423-2446): report the stack and stop with a failure
                std::printf("BREAKPOINT INTERRUPT ISSUED\nStack trace:\n");
                for (size_t k = 0; k < st.size(); ++k)
                    std::printf("    %zu: %s of %zu points\n", k, st[k].spectrum ? "SPECTRUM" : st[k].sf ? "SF" : "CF",
                                st[k].x.size());
                rc = fail(HH_E_INVALID, "INT3: breakpoint interrupt at PROCESSING call %d", i + 1);
                break;
            }
            case HH_PROC_CMP: {
                if ((rc = need(2)) < 0) break;
                std::string why;
                const bool same = cmp_series(st[st.size() - 1], st[st.size() - 2], why);
                st.pop_back();
                st.pop_back();
                if (!quiet) std::printf("CMP: %s%s\n", same ? "equal" : "differ: ", same ? "" : why.c_str());
                if (!same) code = fail(HH_E_INVALID, "CMP failed: %s", why.c_str());
                break;
            }
            case HH_PROC_DUP:
                if ((rc = need(1)) == HH_OK) st.push_back(st.back());
                break;
            case HH_PROC_DUP2:
                if ((rc = need(2)) == HH_OK) {
                    Series a = st[st.size() - 2], b = st[st.size() - 1];
                    st.push_back(a);
                    st.push_back(b);
                }
                break;
            case HH_PROC_DROP:
                if ((rc = need(1)) == HH_OK) st.pop_back();
                break;
            case HH_PROC_DROP2:
                if ((rc = need(2)) == HH_OK) st.resize(st.size() - 2);
                break;
            case HH_PROC_READ_SPECTRUM: {
                Series s;
                rc = read_series(p.arg, false, s);
                s.spectrum = true;
                if (rc == HH_OK) st.push_back(std::move(s));
                break;
            }
            case HH_PROC_D1:
            case HH_PROC_D2:
            case HH_PROC_D3:
            case HH_PROC_ALPHA: {
                if ((rc = need(1)) < 0) break;
                if (!st.back().sf || st.back().spectrum) {
                    rc = fail(HH_E_INVALID, "PROCESSING call %d expects a spectral function on the stack", i + 1);
                    break;
                }
                if (!(st.back().T > 0)) {
                    rc = fail(HH_E_INVALID, "PROCESSING call %d: the spectral function has no temperature", i + 1);
                    break;
                }
                st.back() = p.op == HH_PROC_ALPHA ? alpha(st.back())
                                                  : desymmetrize(st.back(), p.op == HH_PROC_D1 ? 1 : p.op == HH_PROC_D2 ? 2 : 3);
                break;
            }
            case HH_PROC_D4:
            case HH_PROC_D4A: {
                if ((rc = need(1)) < 0) break;
                if (st.back().sf || st.back().spectrum) {
                    rc = fail(HH_E_INVALID, "D4/D4a expect a correlation function on the stack");
                    break;
                }
                Series sf;
                rc = egelstaff_sf(e, st.back(), p.op == HH_PROC_D4A, sf);
                if (rc == HH_OK) st.back() = std::move(sf);
                break;
            }
            case HH_PROC_AVERAGE_CFS: {
                // every CF on the stack, weighted by its trajectory count (average_correlation_functions,
                // src/hawaii.c:4599-4673); the average replaces the whole stack
                if ((rc = need(1)) < 0) break;
                Series avg;
                double wsum = 0.0;
                for (size_t k = 0; k < st.size() && rc == HH_OK; ++k) {
                    const Series& cf = st[k];
                    if (cf.sf || cf.spectrum) {
                        rc = fail(HH_E_INVALID, "AVERAGE_CFS expects only correlation functions on the stack");
                    } else if (!(cf.ntraj > 0)) {
                        rc = fail(HH_E_INVALID, "# of trajectories for correlation function (%zu) is invalid: %.4e", k,
                                  cf.ntraj);
                    } else if (k == 0) {
                        if (cf.x.size() < 2 || !(cf.x[1] - cf.x[0] > 0.0) || !(cf.T > 0))
                            rc = fail(HH_E_INVALID, "AVERAGE_CFS: the first CF needs two samples and a temperature");
                        avg = cf;
                        for (double& v : avg.y) v = 0.0;
                    } else if (cf.x.size() != avg.x.size()) {
                        rc = fail(HH_E_INVALID, "expected correlation functions to be of equal length, however cf.len = "
                                                "%zu and expected len = %zu", cf.x.size(), avg.x.size());
                    } else if (std::fabs(cf.T - avg.T) > 1e-10) {
                        rc = fail(HH_E_INVALID, "AVERAGE_CFS: mismatching temperatures %.2e and %.2e", cf.T, avg.T);
                    } else if (std::fabs((cf.x[1] - cf.x[0]) - (avg.x[1] - avg.x[0])) > 1e-12) {
                        rc = fail(HH_E_INVALID, "AVERAGE_CFS: the time steps of the correlation functions differ");
                    }
                    if (rc != HH_OK) break;
                    for (size_t j = 0; j < avg.y.size(); ++j) avg.y[j] += cf.ntraj * cf.y[j];
                    wsum += cf.ntraj;
                }
                if (rc != HH_OK) break;
                for (double& v : avg.y) v /= wsum;
                avg.ntraj = wsum;
                st.clear();
                st.push_back(std::move(avg));
                break;
            }
            case HH_PROC_MN_CLASSICAL:
            case HH_PROC_MN_QUANTUM: {
                // execute_compute_mn_{classical,quantum}_detailed_balance (src/driver.c:2304-2421):
                // consumes the CF or SF on top; an SF is cut at SPECTRUM_FREQUENCY_MAX first
                if ((rc = need(1)) < 0) break;
                const int n = std::atoi(p.arg);
                Series it = std::move(st.back());
                st.pop_back();
                const bool quantum = p.op == HH_PROC_MN_QUANTUM;
                if (it.spectrum || (quantum && !it.sf)) {
                    rc = fail(HH_E_INVALID, "spectral moments are computed from %s", quantum ? "an SF" : "a CF or an SF");
                    break;
                }
                double Mn = 0.0;
                if (it.sf) {
                    if (c->spectrum_frequency_max > 0 && it.x.size() > 1) {
                        const size_t keep = (size_t)(c->spectrum_frequency_max / (it.x[1] - it.x[0]));
                        if (it.x.size() > keep) {
                            it.x.resize(keep);
                            it.y.resize(keep);
                        }
                    }
                    if (quantum) {
                        // compute_Mn_from_sf_using_quantum_detailed_balance (src/hawaii.c:6011-6046)
                        const double coeff = (8.0 * M_PI * M_PI * M_PI) * 2.686780111e25 * 2.686780111e25 / 3.0 / kHBarSI;
                        std::vector<double> y(it.x.size());
                        for (size_t k = 0; k < y.size(); ++k) {
                            const double e = std::exp(-h_nu_over_kT(it.x[k], it.T));
                            y[k] = it.y[k] * std::pow(it.x[k], n) * (n % 2 == 0 ? 1.0 + e : 1.0 - e);
                        }
                        const double h = it.x[1] - it.x[0];
                        double sum = y.front() + y.back();
                        for (size_t k = 1; k + 1 < y.size(); ++k) sum += (k % 3 == 0 ? 2.0 : 3.0) * y[k];
                        Mn = coeff * (sum * 3.0 * h / 8.0);
                    } else {
                        rc = hh_sf_moment(it.x.data(), it.y.data(), (long)it.x.size(), 0.0, n, &Mn);
                    }
                } else {
                    rc = hh_cf_moment(it.y.data(), (long)it.y.size(), it.x.size() > 1 ? it.x[1] - it.x[0] : 0.0, 0.0, n,
                                      &Mn);
                }
                if (rc == HH_OK && !quiet)
                    std::printf("Spectral moment using %s detailed balance [based on %s] = %.5e\n",
                                quantum ? "quantum" : "classical", it.sf ? "SF" : "CF", Mn);
                break;
            }
            case HH_PROC_ADD_SPECTRA: {
                // execute_add_spectra (src/driver.c:2474-2621): the top n spectra (all when n < 0)
                // resampled by natural cubic splines onto 0, step, ... up to the smallest maximum
                // frequency (step: the largest input step unless given) and summed
                long n = -1;
                double step = 0.0;
                std::sscanf(p.arg, "%ld;%lf", &n, &step);
                const bool step_set = step > 0.0;
                const size_t cnt = n < 0 ? st.size() : (size_t)n;
                if (cnt == 0 || cnt > st.size()) {
                    rc = fail(HH_E_INVALID, "requested to add %zu spectra but only %zu found on stack", cnt, st.size());
                    break;
                }
                double fmax = 3.4e38;
                for (size_t k = st.size() - cnt; k < st.size() && rc == HH_OK; ++k) {
                    const Series& sp = st[k];
                    if (!sp.spectrum || sp.x.size() < 2) {
                        rc = fail(HH_E_INVALID, "ADD_SPECTRA expects spectra on the stack");
                        break;
                    }
                    const double s = (sp.x.back() - sp.x.front()) / (double)(sp.x.size() - 1);
                    const double fb = s * (double)(sp.x.size() - 1);
                    if (!step_set && s > step) step = s;
                    if (fmax > fb) fmax = fb;
                    if (k > st.size() - cnt && sp.T != st[k - 1].T)
                        rc = fail(HH_E_INVALID, "temperature mismatch detected at spectrum (index: %zu)", k);
                }
                if (rc != HH_OK) break;
                size_t len = (size_t)std::ceil(fmax / step) + 1;
                while ((double)(len - 1) * step > fmax) --len;
                Series sum;
                sum.spectrum = true;
                sum.T = st.back().T;
                sum.x.resize(len);
                sum.y.assign(len, 0.0);
                for (size_t j = 0; j < len; ++j) sum.x[j] = step * (double)j;   // linspace(0, step (len - 1), len)
                for (size_t k = st.size() - cnt; k < st.size(); ++k) {
                    const NaturalSpline spl(st[k].x, st[k].y);
                    for (size_t j = 0; j < len; ++j) sum.y[j] += spl((double)j * step);
                    sum.ntraj += st[k].ntraj;
                }
                st.resize(st.size() - cnt);
                st.push_back(std::move(sum));
                break;
            }
            case HH_PROC_WRITE_SPECTRUM: {
                if ((rc = need(1)) < 0) break;
                if (!st.back().spectrum) {
                    rc = fail(HH_E_INVALID, "WRITE_SPECTRUM: wrong element type on the stack");
                    break;
                }
                Series sp = st.back();
                // SPECTRUM_FREQUENCY_MAX truncation (execute_write_spectrum, src/driver.c:2831-2838)
                if (c->spectrum_frequency_max > 0 && sp.x.size() > 1) {
                    const size_t keep = (size_t)(c->spectrum_frequency_max / (sp.x[1] - sp.x[0]));
                    if (sp.x.size() > keep) {
                        sp.x.resize(keep);
                        sp.y.resize(keep);
                    }
                }
                rc = write_series(p.arg, sp);
                st.pop_back();
                break;
            }
            case HH_PROC_WRITE_CF:
            case HH_PROC_WRITE_SF:
                if ((rc = need(1)) < 0) break;
                if (st.back().spectrum || st.back().sf != (p.op == HH_PROC_WRITE_SF)) {
                    rc = fail(HH_E_INVALID, "WRITE_%s: wrong element type on the stack", p.op == HH_PROC_WRITE_SF ? "SF" : "CF");
                    break;
                }
                rc = write_series(p.arg, st.back());
                st.pop_back();
                break;
            default:
                rc = fail(HH_E_INVALID, "unknown PROCESSING op %d", p.op);
        }
        if (rc < 0) return rc;
    }
    return code;
}

uint32_t good_seed()
{
    uint32_t s = 0;
    FILE* f = std::fopen("/dev/urandom", "rb");
    if (f) {
        if (std::fread(&s, sizeof s, 1, f) != 1) s = 0;
        std::fclose(f);
    }
    if (s == 0) s = (uint32_t)time(nullptr);
    return s;
}

// the engine's system for the configured monomers (ATOM+ATOM: He-Ar surfaces, LINEAR_MOLECULE+ATOM:
// CO2-Ar surfaces, ROTOR+LINEAR_MOLECULE: CH4-CO2 surfaces)
int system_of(const hh_conf* c, hh_system* sys)
{
    *sys = hh_system{};
    sys->mu = c->reduced_mass;
    const int t0 = c->monomer_type[0], t1 = c->monomer_type[1];
    if (c->n_monomers == 2 && t0 == HH_MONOMER_ATOM && t1 == HH_MONOMER_ATOM) {
        sys->system = HH_SYS_HEAR;
    } else if (c->n_monomers == 2 && t0 == HH_MONOMER_LINEAR_MOLECULE && t1 == HH_MONOMER_ATOM) {
        sys->system = HH_SYS_CO2AR;
        sys->I = c->II[0][0];
        if (c->II[0][0] != c->II[0][1] || sys->I <= 0)
            return fail(HH_E_INVALID, "linear molecule needs II = {I, I, 0} with I > 0");
    } else if (c->n_monomers == 2 && t0 == HH_MONOMER_ROTOR && t1 == HH_MONOMER_LINEAR_MOLECULE) {
        sys->system = HH_SYS_CH4CO2;
        for (int k = 0; k < 3; ++k) sys->I_rotor[k] = c->II[0][k];
        sys->I = c->II[1][0];
        if (!(sys->I_rotor[0] > 0 && sys->I_rotor[1] > 0 && sys->I_rotor[2] > 0))
            return fail(HH_E_INVALID, "rotor needs three positive principal moments II");
        if (c->II[1][0] != c->II[1][1] || sys->I <= 0)
            return fail(HH_E_INVALID, "linear molecule needs II = {I, I, 0} with I > 0");
    } else {
        return fail(HH_E_UNSUPPORTED, "monomers other than ATOM+ATOM, LINEAR_MOLECULE+ATOM or "
                                      "ROTOR+LINEAR_MOLECULE are not part of this engine");
    }
    if (!(sys->mu > 0)) return fail(HH_E_INVALID, "pair_reduced_mass must be positive");
    return HH_OK;
}

// The plugins named by so_potential / so_dipole (src/driver.c:3083-3181): pes_lab and dpes_lab,
// dipole_lab, their optional pes_init / dipole_init(bool) run once at load.  A plugin that cannot
// be loaded or lacks a required symbol is an error, as in the reference's loader (which exits).
int load_surfaces(const hh_conf* c, HostSurfaces& hs)
{
    auto declared = [](void* so) {
        auto id = reinterpret_cast<int (*)(void)>(dlsym(so, "hh_system_id"));
        return id ? id() : -1;
    };
    if (c->so_potential[0]) {
        void* so = dlopen(c->so_potential, RTLD_LAZY);
        if (!so) return fail(HH_E_INVALID, "cannot load the surface plugin %s: %s", c->so_potential, dlerror());
        hs.so[0] = so;
        hs.pes = reinterpret_cast<hh_pes_fn>(dlsym(so, "pes_lab"));
        hs.dpes = reinterpret_cast<hh_dpes_fn>(dlsym(so, "dpes_lab"));
        if (!hs.pes || !hs.dpes) return fail(HH_E_INVALID, "%s must export pes_lab and dpes_lab", c->so_potential);
        hs.pes_system = declared(so);
        hs.pes_name = c->so_potential;
        if (void (*init)(void) = reinterpret_cast<void (*)(void)>(dlsym(so, "pes_init"))) init();
    }
    if (c->so_dipole[0]) {
        void* so = dlopen(c->so_dipole, RTLD_LAZY);
        if (!so) return fail(HH_E_INVALID, "cannot load the surface plugin %s: %s", c->so_dipole, dlerror());
        hs.so[1] = so;
        hs.dipole = reinterpret_cast<hh_dipole_fn>(dlsym(so, "dipole_lab"));
        if (!hs.dipole) return fail(HH_E_INVALID, "%s must export dipole_lab", c->so_dipole);
        hs.dipole_system = declared(so);
        hs.dipole_name = c->so_dipole;
        if (void (*init)(bool) = reinterpret_cast<void (*)(bool)>(dlsym(so, "dipole_init"))) init(false);
    }
    return HH_OK;
}

// The host surfaces must be the system the device integrates: the engine evaluates only its
// built-in surfaces, so a surface is accepted when its plugin declares the same system
// (`int hh_system_id(void)` returning the HH_SYS_* value) or when pes_lab / dpes_lab /
// dipole_lab agree with the built-in surface at 48 fixed points of [Rmin, min(Rmax, 20)] and
// the angles' ranges (relative 1e-6 of the largest magnitude, FP64 device evaluation).  Absent
// surfaces (null pointers) are the built-in ones.
int check_surfaces(hh_engine* e, const hh_system& sys, double Rmin, double Rmax, const HostSurfaces& hs)
{
    const char* names[] = {"He-Ar", "CO2-Ar", "N2-Ar", "CH4-CO2"};
    const char* name = names[sys.system];
    for (int w = 0; w < 2; ++w) {
        const int sid = w ? hs.dipole_system : hs.pes_system;
        if (sid >= 0 && sid != sys.system)
            return fail(HH_E_UNSUPPORTED, "%s declares system %d; the configured monomers are the engine's %s",
                        (w ? hs.dipole_name : hs.pes_name).c_str(), sid, name);
    }
    const bool check_pes = hs.pes && hs.pes_system < 0, check_dip = hs.dipole && hs.dipole_system < 0;
    if (!check_pes && !check_dip) return HH_OK;
    const int qp = hh_qp_size(sys.system), nq = qp / 2;
    constexpr int K = 48;
    // points: R over [Rmin, min(Rmax, 20)] (the interaction region), angles over their ranges,
    // from a fixed linear congruential stream; momenta zero (surfaces depend on coordinates only)
    const double r0 = Rmin > 0 ? Rmin : 4.0;
    const double r1 = std::max(r0 + 1.0, std::min(Rmax > 0 ? Rmax : 20.0, 20.0));
    std::vector<double> y((size_t)K * qp, 0.0), q((size_t)K * nq);
    uint64_t st = 0x9e3779b97f4a7c15ull;
    auto u = [&] {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        return (double)(st >> 11) / 9007199254740992.0;
    };
    for (int k = 0; k < K; ++k)
        for (int i = 0; i < nq; ++i) {
            // coordinate order Phi, Theta, R, then the monomer angles (phi, theta[, psi])
            double v;
            if (i == 2)
                v = r0 + (r1 - r0) * u();
            else if (i == 1 || (i >= 3 && (i - 3) % 2 == 1))
                v = std::acos(2.0 * u() - 1.0);
            else
                v = 2.0 * M_PI * u();
            q[(size_t)k * nq + i] = v;
            y[(size_t)(2 * i) * K + k] = v;   // SoA [QP][K], coordinate i at state index 2i
        }
    double *y_dev = nullptr, *out_dev = nullptr;
    if (hipMalloc(&y_dev, y.size() * sizeof(double)) != hipSuccess ||
        hipMalloc(&out_dev, (size_t)qp * K * sizeof(double)) != hipSuccess) {
        if (y_dev) (void)hipFree(y_dev);
        return fail(HH_E_NOMEM, "surface check: device allocation failed");
    }
    struct Free {
        double *a, *b;
        ~Free()
        {
            (void)hipFree(a);
            (void)hipFree(b);
        }
    } fr{y_dev, out_dev};
    HH_CHECK(hipMemcpy(y_dev, y.data(), y.size() * sizeof(double), hipMemcpyHostToDevice));
    auto device = [&](int what, int rows, std::vector<double>& out) -> int {
        const int rc = hh_eval(e, &sys, what, y_dev, K, out_dev);
        if (rc < 0) return rc;
        out.resize((size_t)rows * K);
        HH_CHECK(hipMemcpy(out.data(), out_dev, out.size() * sizeof(double), hipMemcpyDeviceToHost));
        return HH_OK;
    };
    if (check_pes) {
        std::vector<double> V, F;
        int rc;
        if ((rc = device(HH_EVAL_PES, 1, V)) < 0 || (rc = device(HH_EVAL_RHS, qp, F)) < 0) return rc;
        double scale = 0.0, worst = 0.0, worst_g = 0.0, gscale = 0.0;
        std::vector<double> qq(nq), g(nq);
        for (int k = 0; k < K; ++k) {
            for (int i = 0; i < nq; ++i) qq[i] = q[(size_t)k * nq + i];
            const double vp = hs.pes(qq.data());
            scale = std::max(scale, std::fabs(V[k]));
            worst = std::max(worst, std::fabs(vp - V[k]));
            if (!hs.dpes) continue;
            hs.dpes(qq.data(), g.data());
            // the momentum derivatives at zero momenta are -dV/dq
            for (int i = 0; i < nq; ++i) {
                const double gd = -F[(size_t)(2 * i + 1) * K + k];
                gscale = std::max(gscale, std::fabs(gd));
                worst_g = std::max(worst_g, std::fabs(g[i] - gd));
            }
        }
        if (!(worst <= 1e-6 * scale) || !(worst_g <= 1e-6 * gscale))
            return fail(HH_E_UNSUPPORTED, "%s is not the engine's built-in %s surface (max |dV| %.3e of %.3e Hartree, "
                                          "max |d grad| %.3e of %.3e): the device integrates only its built-in surfaces",
                        hs.pes_name.c_str(), name, worst, scale, worst_g, gscale);
    }
    if (check_dip) {
        std::vector<double> M;
        const int rc = device(HH_EVAL_DIPOLE, 3, M);
        if (rc < 0) return rc;
        double scale = 0.0, worst = 0.0;
        std::vector<double> qq(nq);
        for (int k = 0; k < K; ++k) {
            for (int i = 0; i < nq; ++i) qq[i] = q[(size_t)k * nq + i];
            double mu[3];
            hs.dipole(qq.data(), mu);
            for (int j = 0; j < 3; ++j) {
                scale = std::max(scale, std::fabs(M[(size_t)j * K + k]));
                worst = std::max(worst, std::fabs(mu[j] - M[(size_t)j * K + k]));
            }
        }
        if (!(worst <= 1e-6 * scale))
            return fail(HH_E_UNSUPPORTED, "%s is not the engine's built-in %s dipole (max |d mu| %.3e of %.3e a.u.): "
                                          "the device evaluates only its built-in surfaces",
                        hs.dipole_name.c_str(), name, worst, scale);
    }
    return HH_OK;
}

// the engine's system for the configured monomers, checked against the configured plugins
int resolve_system(hh_engine* e, const hh_conf* c, const HostSurfaces& hs, hh_system* sys)
{
    const int rc = system_of(c, sys);
    if (rc < 0) return rc;
    return check_surfaces(e, *sys, c->sampler_Rmin, c->sampler_Rmax, hs);
}

// The M2 estimates from the accumulated CF after each iteration (src/hawaii.c:3070-3090,
// 2735-2751): the 9-point stencil over the first 5 samples, the 21-point one over 11, each
// beside the VEGAS M2 (0 where none was integrated).
void print_m2_from_cf(const Series& cf, double hep, const char* prefix)
{
    const long len = (long)cf.y.size();
    const double dt = cf.x.size() > 1 ? cf.x[1] - cf.x[0] : 0.0;
    double m2 = 0.0;
    if (len >= 5 && hh_cf_moment(cf.y.data(), 5, dt, 0.0, 2, &m2) == HH_OK)
        std::printf("%sM2 ESTIMATE FROM CF (9-point): %.5e, PRELIMINARY M2 ESTIMATE: %.5e, diff: %.3f%%\n", prefix, m2,
                    hep, (m2 - hep) / hep * 100.0);
    else
        std::printf("%sTrajectory is too short to estimate M2\n", prefix);
    if (len >= 11 && hh_cf_moment(cf.y.data(), len, dt, 0.0, 2, &m2) == HH_OK)
        std::printf("%sM2 ESTIMATE FROM CF (21-point): %.5e, PRELIMINARY M2 ESTIMATE: %.5e, diff: %.3f%%\n", prefix,
                    m2, hep, (m2 - hep) / hep * 100.0);
}

// CALCULATION_PHASE_SPACE_M0 (src/driver.c:3389-3445): VEGAS M0 at each temperature
int run_phase_space_m0(hh_engine* e, const hh_conf* c, const hh_conf_run_opts& o, const HostSurfaces& hs)
{
    if (!(c->sampler_Rmin > 0) || !(c->sampler_Rmax > c->sampler_Rmin))
        return fail(HH_E_INVALID, "sampler_Rmin and sampler_Rmax must be set for the phase-space integral");
    hh_system sys{};
    int rc = resolve_system(e, c, hs, &sys);
    if (rc < 0) return rc;
    std::vector<double> Ts(c->temperatures, c->temperatures + c->ntemperatures);
    if (Ts.empty() && c->temperature > 0) Ts.push_back(c->temperature);
    if (Ts.empty()) return fail(HH_E_INVALID, "no temperature is provided to run CALCULATION_PHASE_SPACE_M0");
    for (double T : Ts) {
        const double pf_an = hh_pf_analytic(&sys, T);
        if (pf_an < 0) return (int)HH_E_INVALID;
        if (!o.quiet) std::printf("Analytic partition function divided by V: %.5e\n", pf_an);
        hh_vegas_params vp{HH_INTEGRAND_M0, c->pair_state ? c->pair_state : HH_PAIR_FREE, T, c->sampler_Rmin,
                           c->sampler_Rmax, (int)(c->hep_m0_niterations > 0 ? c->hep_m0_niterations : 12),
                           c->hep_m0_npoints > 0 ? c->hep_m0_npoints : 1000000L, o.seed ? o.seed : 1u};
        hh_vegas_result vr{};
        rc = hh_vegas(e, &sys, &vp, &vr);
        if (rc < 0) return rc;
        std::printf("T = %.2e => M0: %.5e +- %.5e\n", T, vr.value * kZeroCoeff / pf_an, vr.error * kZeroCoeff / pf_an);
    }
    return HH_OK;
}

// The VEGAS M2 at T (c_mpi_perform_integration(INTEGRAND_M2), scaled by SecondCoeff / pf_analytic
// as src/driver.c:3377-3381 and src/hawaii.c:2867-2880 do), with HEP_M2_NITERATIONS x
// HEP_M2_NPOINTS (default 12 x 10^6).
int hep_m2(hh_engine* e, const hh_system& sys, const hh_conf* c, int pair_state, double T, uint32_t seed, double* m2,
           double* m2_err)
{
    const double pf_an = hh_pf_analytic(&sys, T);
    if (pf_an < 0) return (int)HH_E_INVALID;
    hh_vegas_params vp{HH_INTEGRAND_M2, pair_state, T, c->sampler_Rmin, c->sampler_Rmax,
                       (int)(c->hep_m2_niterations > 0 ? c->hep_m2_niterations : 12),
                       c->hep_m2_npoints > 0 ? c->hep_m2_npoints : 1000000L, seed};
    hh_vegas_result vr{};
    const int rc = hh_vegas(e, &sys, &vp, &vr);
    if (rc < 0) return rc;
    *m2 = vr.value * kSecondCoeff / pf_an;
    if (m2_err) *m2_err = vr.error * kSecondCoeff / pf_an;
    return HH_OK;
}

// CALCULATION_PHASE_SPACE_M2 (src/driver.c:3344-3388): VEGAS M2 at TEMPERATURE
int run_phase_space_m2(hh_engine* e, const hh_conf* c, const hh_conf_run_opts& o, const HostSurfaces& hs)
{
    if (!(c->sampler_Rmin > 0))
        return fail(HH_E_INVALID, "minimum value for R throught integration (sampler_Rmin) must be set");
    if (!(c->sampler_Rmax > 0))
        return fail(HH_E_INVALID, "maximum value for R throught integration (sampler_Rmax) must be set");
    if (!(c->temperature > 0)) return fail(HH_E_INVALID, "no temperature is provided to run CALCULATION_PHASE_SPACE_M2");
    hh_system sys{};
    int rc = resolve_system(e, c, hs, &sys);
    if (rc < 0) return rc;
    const double pf_an = hh_pf_analytic(&sys, c->temperature);
    if (pf_an < 0) return (int)HH_E_INVALID;
    if (!o.quiet) std::printf("Analytic partition function divided by V: %.5e\n", pf_an);
    double m2 = 0.0, m2_err = 0.0;
    rc = hep_m2(e, sys, c, c->pair_state ? c->pair_state : HH_PAIR_FREE, c->temperature, o.seed ? o.seed : 1u, &m2,
                &m2_err);
    if (rc < 0) return rc;
    std::printf("T = %.2e => M2: %.5e +- %.5e\n", c->temperature, m2, m2_err);
    return HH_OK;
}

int run_correlation_single(hh_engine* e, const hh_conf* c, const hh_conf_run_opts& o, const HostSurfaces& hs,
                           RunIO* io)
{
    // the reference's checks (src/hawaii.c:2802-2811)
    if (c->max_trajectory_length <= 0 || c->Rcut <= 0 || c->sampling_time <= 0 || c->total_trajectories <= 0 ||
        c->cvode_tolerance <= 0 || c->niterations < 1 || !c->cf_filename[0])
        return fail(HH_E_INVALID, "MaxTrajectoryLength, Rcut, sampling_time, total_trajectories, cvode_tolerance, "
                                  "niterations and cf_filename must be set");
    // src/hawaii.c:2942-2945
    if (c->use_zimmermann_trick && c->pair_state != HH_PAIR_BOUND)
        return fail(HH_E_INVALID, "Zimmermann's trick can be used only for bound states");
    if (c->pesmin == 0 || !(c->sampler_Rmin > 0 && c->sampler_Rmax > c->sampler_Rmin))
        return fail(HH_E_INVALID, "pesmin, sampler_Rmin and sampler_Rmax must be set");
    hh_system sys{};
    {
        const int rc0 = resolve_system(e, c, hs, &sys);
        if (rc0 < 0) return rc0;
    }
    // partition-function ratio, M0 and M2 by VEGAS when the ratio is not given (src/hawaii.c:2829-2880);
    // otherwise the M2 the CF estimates are compared with stays 0 (src/hawaii.c:2867)
    double ppf = c->ppf;
    double hep_m0 = 0.0, prelim_m2 = 0.0;
    if (ppf <= 0) {
        const double pf_an = hh_pf_analytic(&sys, c->temperature);
        if (pf_an < 0) return (int)HH_E_INVALID;
        if (!o.quiet) {
            std::printf("Ratio of partial partition functions to full (analytic) partial partition function is not "
                        "provided.\nConducting calculations using adaptive Monte Carlo method.\n\n\n");
            std::printf("Analytic partition function divided by V: %.5e\n", pf_an);
        }
        hh_vegas_params vp{HH_INTEGRAND_PF, c->pair_state ? c->pair_state : HH_PAIR_FREE, c->temperature,
                           c->sampler_Rmin, c->sampler_Rmax,
                           (int)(c->hep_ppf_niterations > 0 ? c->hep_ppf_niterations : 12),
                           c->hep_ppf_npoints > 0 ? c->hep_ppf_npoints : 1000000L, o.seed ? o.seed : 1u};
        hh_vegas_result vr{};
        int vrc = hh_vegas(e, &sys, &vp, &vr);
        if (vrc < 0) return vrc;
        ppf = vr.value / pf_an;
        if (!o.quiet) std::printf("T = %.2e => PPF ratio: %.5e\n", c->temperature, ppf);
        vp.integrand = HH_INTEGRAND_M0;
        vp.niterations = (int)(c->hep_m0_niterations > 0 ? c->hep_m0_niterations : 12);
        vp.npoints = c->hep_m0_npoints > 0 ? c->hep_m0_npoints : 1000000L;
        vrc = hh_vegas(e, &sys, &vp, &vr);
        if (vrc < 0) return vrc;
        hep_m0 = vr.value * kZeroCoeff / pf_an;
        if (!o.quiet) std::printf("T = %.2e => M0: %.5e\n", c->temperature, hep_m0);
        vrc = hep_m2(e, sys, c, vp.pair_state, c->temperature, vp.seed, &prelim_m2, nullptr);
        if (vrc < 0) return vrc;
        if (!o.quiet) std::printf("T = %.2e => M2: %.5e\n", c->temperature, prelim_m2);
    }

    const int qp = hh_qp_size(sys.system);
    const long n_iter = c->total_trajectories / c->niterations;
    if (n_iter <= 0) return fail(HH_E_INVALID, "total_trajectories < niterations");
    const uint32_t seed = o.seed ? o.seed : good_seed();

    // the plugin's pes_lab drives the host rejection sampler (the reference's accept/reject)
    hh_sampler smp{HH_IC_REJECTION, c->pair_state ? c->pair_state : HH_PAIR_FREE, c->temperature, c->sampler_Rmin,
                   c->sampler_Rmax, c->pesmin, 0.0};
    // the stream: initialM0_npoints accepted points for the preliminary M0, then
    // initialM2_npoints for the preliminary M2, then the trajectories (src/hawaii.c:2947-2965)
    const long m0n = c->initialM0_npoints > 1 ? c->initialM0_npoints : 0;
    const long m2n = c->initialM2_npoints > 1 ? c->initialM2_npoints : 0;
    const long first_traj = m0n + m2n;
    const long total = first_traj + n_iter * c->niterations;
    std::vector<double> ic((size_t)total * qp);
    long nd = 0;
    if (!o.quiet) std::printf("Generator seed is set to %u\n", seed);
    long rc = hh_sample_ics_from(&sys, &smp, seed, io ? io->skip_draws : 0, total, hs.pes, ic.data(), &nd);
    if (rc < 0) return (int)rc;
    if (io) io->draws_used = nd;

    // device buffers: one iteration of initial conditions (SoA), the CF, the M0 dipoles
    const long nmax = std::max(n_iter, std::max(m0n, m2n));
    double *ic_dev = nullptr, *cf_dev = nullptr, *out_dev = nullptr;
    int code = HH_OK;
    std::vector<double> soa((size_t)nmax * qp);
    auto upload = [&](long first, long n) -> int {
        for (long t = 0; t < n; ++t)
            for (int i = 0; i < qp; ++i) soa[(size_t)i * n + t] = ic[(size_t)(first + t) * qp + i];
        if (hipMemcpy(ic_dev, soa.data(), (size_t)n * qp * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
            return fail(HH_E_DEVICE, "upload of initial conditions failed");
        return (int)HH_OK;
    };
    do {
        if (hipMalloc(&ic_dev, (size_t)nmax * qp * sizeof(double)) != hipSuccess ||
            hipMalloc(&cf_dev, c->max_trajectory_length * sizeof(double)) != hipSuccess ||
            hipMalloc(&out_dev, (size_t)3 * std::max(std::max(m0n, m2n), 1L) * sizeof(double)) != hipSuccess) {
            code = fail(HH_E_NOMEM, "device allocation failed");
            break;
        }
        // preliminary M0 over the first initialM0_npoints accepted points (calculate_M0,
        // src/hawaii.c:1423-1491: the same candidates the CF stream then continues after)
        double m0 = 0.0;
        if (m0n > 0) {
            if ((code = upload(0, m0n)) < 0) break;
            if ((code = hh_eval(e, &sys, HH_EVAL_DIPOLE, ic_dev, m0n, out_dev)) < 0) break;
            std::vector<double> d((size_t)3 * m0n);
            if (hipMemcpy(d.data(), out_dev, d.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
                code = fail(HH_E_DEVICE, "download failed");
                break;
            }
            double m = 0.0, q = 0.0;
            for (long k = 0; k < m0n; ++k) {
                const double f = d[k] * d[k] + d[m0n + k] * d[m0n + k] + d[2 * m0n + k] * d[2 * m0n + k];
                const double diff = f - m;
                m += diff / (k + 1.0);
                q += diff * diff * (k / (k + 1.0));
            }
            m0 = m * kZeroCoeff * ppf;
            const double m0s = std::sqrt(q / m0n / (m0n - 1.0)) * kZeroCoeff * ppf;
            if (!o.quiet) {
                std::printf("M0 = %.10e +/- %.10e [%.10e ... %.10e]\n", m0, m0s, m0 - m0s, m0 + m0s);
                std::printf("Error: %.3f%%\n", m0s / m0 * 100.0);
            }
        }
        // preliminary M2 over the next initialM2_npoints accepted points (calculate_M2,
        // src/hawaii.c:1524-1617): |d mu/dt|^2 on the device, the running mean on the host
        if (m2n > 0) {
            if ((code = upload(m0n, m2n)) < 0) break;
            if ((code = hh_eval(e, &sys, HH_EVAL_DIPOLE_RATE, ic_dev, m2n, out_dev)) < 0) break;
            std::vector<double> f((size_t)m2n);
            if (hipMemcpy(f.data(), out_dev, f.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
                code = fail(HH_E_DEVICE, "download failed");
                break;
            }
            double m = 0.0, q = 0.0;
            for (long k = 0; k < m2n; ++k) {
                const double diff = f[k] - m;
                m += diff / (k + 1.0);
                q += diff * diff * (k / (k + 1.0));
            }
            const double m2 = m * kSecondCoeff * ppf;
            const double m2s = std::sqrt(q / m2n / (m2n - 1.0)) * kSecondCoeff * ppf;
            if (!o.quiet) {
                std::printf("M2 = %.10e +/- %.10e [%.10e ... %.10e]\n", m2, m2s, m2 - m2s, m2 + m2s);
                std::printf("Error: %.3f%%\n", m2s / m2 * 100.0);
            }
        }
        hh_cf_params p{};
        const bool mixed = o.precision == HH_PREC_MIXED;
        p.rtol = p.atol = mixed ? (o.rtol_mixed > 0 ? o.rtol_mixed : 1e-10) : c->cvode_tolerance;
        p.orbit_rtol = c->cvode_tolerance;   // orbiting pairs: FP64 at the configured tolerance
        p.dt = c->sampling_time;
        p.Rcut = c->Rcut;
        p.ppf = ppf;
        p.L = (int)c->max_trajectory_length;
        p.halves = HH_HALVES_BOTH;
        p.zimmermann = c->use_zimmermann_trick ? 1 : 0;
        p.max_steps = 50000;
        p.engine = HH_ENGINE_BATCHED;
        p.precision = mixed ? HH_PREC_MIXED : HH_PREC_FP64;
        Series cf;
        cf.T = c->temperature;
        cf.x.resize(p.L);
        cf.y.assign(p.L, 0.0);
        for (int k = 0; k < p.L; ++k) cf.x[k] = c->sampling_time * k;
        std::vector<double> sum(p.L, 0.0), it(p.L);
        double ntraj = 0.0;
        for (long iter = 0; iter < c->niterations && code == HH_OK; ++iter) {
            if ((code = upload(first_traj + iter * n_iter, n_iter)) < 0) break;
            const double* icp[1] = {ic_dev};
            double* cfp[1] = {cf_dev};
            const long nt[1] = {n_iter};
            hh_cf_stats st;
            if ((code = hh_cf_run(e, &sys, &p, 1, icp, nt, cfp, &st)) < 0) break;
            if (hipMemcpy(it.data(), cf_dev, p.L * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
                code = fail(HH_E_DEVICE, "download failed");
                break;
            }
            for (int k = 0; k < p.L; ++k) sum[k] += it[k] * (double)n_iter;
            ntraj += (double)n_iter;
            cf.ntraj = ntraj;
            for (int k = 0; k < p.L; ++k) cf.y[k] = sum[k] / ntraj;
            if (!o.quiet) {
                std::printf("ITERATION %ld/%ld: accumulated %.0f trajectories. Saving the temporary result to '%s'\n",
                            iter + 1, c->niterations, ntraj, c->cf_filename);
                const double m0cf = cf.y[0] * kZeroCoeff / kALU / kALU / kALU;
                std::printf("M0 ESTIMATE FROM CF: %.5e, PRELIMINARY M0 ESTIMATE: %.5e, diff: %.3f%%\n", m0cf, hep_m0,
                            (m0cf - hep_m0) / hep_m0 * 100.0);
                print_m2_from_cf(cf, prelim_m2, "");
                std::printf("device: %.1f ms integration, %lld RHS, %lld failed, %lld step-limited items\n",
                            st.integrate_ms, st.nrhs, st.n_failed, st.n_stepfail);
                std::fflush(stdout);
            }
            if ((code = write_series(c->cf_filename, cf)) < 0) break;
        }
        if (code == HH_OK && c->sf_filename[0]) {
            Series sf;
            if ((code = cf_to_sf(e, cf, sf)) == HH_OK) code = write_series(c->sf_filename, sf);
        }
        if (code == HH_OK && io && io->cf) *io->cf = cf;
    } while (false);
    if (ic_dev) (void)hipFree(ic_dev);
    if (cf_dev) (void)hipFree(cf_dev);
    if (out_dev) (void)hipFree(out_dev);
    return code;
}

// CALCULATION_TYPE = CORRELATION_ARRAY: calculate_correlation_array_and_save
// (src/hawaii.c:2361-2793).  One ensemble sampled at TEMPERATURE (the base temperature) per
// iteration, integrated once and accumulated at every satellite temperature with its Boltzmann
// weight (hh_cf_array); after each iteration every CF file is rewritten with its weight sum as
// the trajectory count.
int run_correlation_array(hh_engine* e, const hh_conf* c, const hh_conf_run_opts& o, const HostSurfaces& hs,
                          RunIO* io)
{
    // the reference's checks (src/hawaii.c:2367-2378)
    if (c->max_trajectory_length <= 0 || c->Rcut <= 0 || c->sampling_time <= 0 || c->total_trajectories <= 0 ||
        c->cvode_tolerance <= 0 || c->niterations < 1 || c->nsatellite < 1 || c->ncf_filenames != c->nsatellite)
        return fail(HH_E_INVALID, "MaxTrajectoryLength, Rcut, sampling_time, total_trajectories, cvode_tolerance, "
                                  "niterations, satellite_temperatures and one cf_filename per satellite "
                                  "temperature must be set");
    if (c->nppf_ratios != 0 && c->nppf_ratios != c->nsatellite)
        return fail(HH_E_INVALID, "partial_partition_function_ratios needs one ratio per satellite temperature");
    if (c->use_zimmermann_trick)
        return fail(HH_E_INVALID, "Zimmermann's trick is not part of this engine's CORRELATION_ARRAY");
    const int ps = c->pair_state ? c->pair_state : HH_PAIR_FREE;
    if (ps != HH_PAIR_FREE && ps != HH_PAIR_BOUND)
        return fail(HH_E_INVALID, "CORRELATION_ARRAY needs PAIR_STATE FREE_AND_METASTABLE or BOUND");
    if (c->pesmin == 0 || !(c->sampler_Rmin > 0 && c->sampler_Rmax > c->sampler_Rmin))
        return fail(HH_E_INVALID, "pesmin, sampler_Rmin and sampler_Rmax must be set");
    hh_system sys{};
    {
        const int rc0 = resolve_system(e, c, hs, &sys);
        if (rc0 < 0) return rc0;
    }
    const int M = c->nsatellite;
    const double Tb = c->temperature;
    // partition-function ratios (and M0) by VEGAS at each satellite temperature when not given
    // (src/hawaii.c:2388-2437)
    std::vector<double> ratios(M, 0.0), hep_m0(M, 0.0), hep_m2s(M, 0.0);
    if (c->nppf_ratios == M) {
        for (int st = 0; st < M; ++st) ratios[st] = c->ppf_ratios[st];
    } else {
        if (!o.quiet)
            std::printf("Ratios of partial partition functions to full (analytic) partial partition function are not "
                        "provided.\nConducting calculations using adaptive Monte Carlo method.\n");
        for (int st = 0; st < M; ++st) {
            const double T = c->satellite_temperatures[st];
            const double pf_an = hh_pf_analytic(&sys, T);
            if (pf_an < 0) return (int)HH_E_INVALID;
            hh_vegas_params vp{HH_INTEGRAND_PF, ps, T, c->sampler_Rmin, c->sampler_Rmax,
                               (int)(c->hep_ppf_niterations > 0 ? c->hep_ppf_niterations : 12),
                               c->hep_ppf_npoints > 0 ? c->hep_ppf_npoints : 1000000L, (o.seed ? o.seed : 1u) + st};
            hh_vegas_result vr{};
            int vrc = hh_vegas(e, &sys, &vp, &vr);
            if (vrc < 0) return vrc;
            ratios[st] = vr.value / pf_an;
            vp.integrand = HH_INTEGRAND_M0;
            vp.niterations = (int)(c->hep_m0_niterations > 0 ? c->hep_m0_niterations : 12);
            vp.npoints = c->hep_m0_npoints > 0 ? c->hep_m0_npoints : 1000000L;
            vrc = hh_vegas(e, &sys, &vp, &vr);
            if (vrc < 0) return vrc;
            hep_m0[st] = vr.value * kZeroCoeff / pf_an;
            vrc = hep_m2(e, sys, c, ps, T, vp.seed, &hep_m2s[st], nullptr);
            if (vrc < 0) return vrc;
            if (!o.quiet) {
                std::printf("Satellite temperature = %.2e\n", T);
                std::printf("Analytic partition function divided by V: %.5e\n", pf_an);
                std::printf("T = %.2e => PPF ratio: %.5e\n", T, ratios[st]);
                std::printf("T = %.2e => M0: %.5e\n", T, hep_m0[st]);
                std::printf("T = %.2e => M2: %.5e\n", T, hep_m2s[st]);
            }
        }
    }

    const int qp = hh_qp_size(sys.system);
    const long n_iter = c->total_trajectories / c->niterations;
    if (n_iter <= 0) return fail(HH_E_INVALID, "total_trajectories < niterations");
    const uint32_t seed = o.seed ? o.seed : good_seed();
    hh_sampler smp{HH_IC_REJECTION, ps, Tb, c->sampler_Rmin, c->sampler_Rmax, c->pesmin, 0.0};
    const long total = n_iter * c->niterations;
    std::vector<double> ic((size_t)total * qp);
    long nd = 0;
    if (!o.quiet) std::printf("Generator seed is set to %u\n", seed);
    long rc = hh_sample_ics_from(&sys, &smp, seed, io ? io->skip_draws : 0, total, hs.pes, ic.data(), &nd);
    if (rc < 0) return (int)rc;
    if (io) io->draws_used = nd;

    hh_cf_params p{};
    const bool mixed = o.precision == HH_PREC_MIXED;
    p.rtol = p.atol = mixed ? (o.rtol_mixed > 0 ? o.rtol_mixed : 1e-10) : c->cvode_tolerance;
    p.orbit_rtol = c->cvode_tolerance;   // orbiting pairs: FP64 at the configured tolerance
    p.dt = c->sampling_time;
    p.Rcut = c->Rcut;
    p.ppf = 1.0;
    p.L = (int)c->max_trajectory_length;
    p.halves = HH_HALVES_BOTH;
    p.max_steps = 50000;
    p.engine = HH_ENGINE_BATCHED;
    p.precision = mixed ? HH_PREC_MIXED : HH_PREC_FP64;
    hh_cf_array_params ap{Tb, ps, M, c->pesmin, c->satellite_temperatures, ratios.data()};

    double *ic_dev = nullptr, *cf_dev = nullptr;
    int code = HH_OK;
    std::vector<double> soa((size_t)n_iter * qp), it((size_t)M * p.L), ns(M);
    std::vector<double> sum((size_t)M * p.L, 0.0), nstar(M, 0.0);
    do {
        if (hipMalloc(&ic_dev, (size_t)n_iter * qp * sizeof(double)) != hipSuccess ||
            hipMalloc(&cf_dev, (size_t)M * p.L * sizeof(double)) != hipSuccess) {
            code = fail(HH_E_NOMEM, "device allocation failed");
            break;
        }
        for (long iter = 0; iter < c->niterations && code == HH_OK; ++iter) {
            const long first = iter * n_iter;
            for (long t = 0; t < n_iter; ++t)
                for (int i = 0; i < qp; ++i) soa[(size_t)i * n_iter + t] = ic[(size_t)(first + t) * qp + i];
            if (hipMemcpy(ic_dev, soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
                code = fail(HH_E_DEVICE, "upload of initial conditions failed");
                break;
            }
            hh_cf_stats st;
            if ((code = hh_cf_array(e, &sys, &p, &ap, ic_dev, n_iter, cf_dev, ns.data(), &st)) < 0) break;
            if (hipMemcpy(it.data(), cf_dev, it.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
                code = fail(HH_E_DEVICE, "download failed");
                break;
            }
            // the reference accumulates unnormalised sums and their weight sums (src/hawaii.c:2681-2697)
            for (int s2 = 0; s2 < M; ++s2) {
                for (int k = 0; k < p.L; ++k) sum[(size_t)s2 * p.L + k] += it[(size_t)s2 * p.L + k] * ns[s2];
                nstar[s2] += ns[s2];
            }
            if (!o.quiet)
                std::printf("\nITERATION %ld/%ld: accumulated %ld trajectories. Saving the temporary results\n",
                            iter + 1, c->niterations, (iter + 1) * n_iter);
            for (int s2 = 0; s2 < M && code == HH_OK; ++s2) {
                Series cf;
                cf.T = c->satellite_temperatures[s2];
                cf.ntraj = nstar[s2];
                cf.x.resize(p.L);
                cf.y.assign(p.L, 0.0);
                for (int k = 0; k < p.L; ++k) {
                    cf.x[k] = c->sampling_time * k;
                    cf.y[k] = nstar[s2] > 0 ? sum[(size_t)s2 * p.L + k] / nstar[s2] : 0.0;
                }
                if (!o.quiet) {
                    const double m0cf = cf.y[0] * kZeroCoeff / kALU / kALU / kALU;
                    std::printf("T = %.2f: NSTAR = %.2f, M0 ESTIMATE FROM CF: %.5e, PRELIMINARY M0 ESTIMATE: %.5e, "
                                "diff: %.3f%%\n", cf.T, nstar[s2], m0cf, hep_m0[s2],
                                (m0cf - hep_m0[s2]) / hep_m0[s2] * 100.0);
                    print_m2_from_cf(cf, hep_m2s[s2], "");
                }
                code = write_series(c->cf_filenames[s2], cf);
                if (io && io->cfs && iter + 1 == c->niterations) io->cfs->push_back(cf);
            }
            if (!o.quiet) std::fflush(stdout);
        }
    } while (false);
    if (ic_dev) (void)hipFree(ic_dev);
    if (cf_dev) (void)hipFree(cf_dev);
    return code;
}

// Welford mean and its standard error over f[0..n) (calculate_M0 / calculate_M2's running sums),
// times coeff
void welford(const std::vector<double>& f, double coeff, double* m_out, double* s_out)
{
    double m = 0.0, q = 0.0;
    const long n = (long)f.size();
    for (long k = 0; k < n; ++k) {
        const double diff = f[k] - m;
        m += diff / (k + 1.0);
        q += diff * diff * (k / (k + 1.0));
    }
    *m_out = m * coeff;
    *s_out = std::sqrt(q / n / (n - 1.0)) * coeff;
}

// CALCULATION_TYPE = PR_MU: calculate_spectral_function_using_prmu_representation_and_save
// (src/hawaii.c:3148-3977) without requantised rotors.  The stream serves the preliminary M0
// (initialM0_npoints) and M2 (initialM2_npoints) first, then each iteration draws R0-injected
// initial conditions until TOTAL_TRAJECTORIES / NITERATIONS of them have ended beyond R0 within
// MaxTrajectoryLength samples (the reference draws a new one in place of each that did not);
// the SF file is rewritten after every iteration.
int run_prmu(hh_engine* e, const hh_conf* c, const hh_conf_run_opts& o, const HostSurfaces& hs, RunIO* io)
{
    // the reference's checks (src/hawaii.c:3160-3168, 1529-1530)
    if (c->max_trajectory_length <= 0 || c->sampling_time <= 0 || c->cvode_tolerance <= 0 ||
        c->total_trajectories <= 0 || c->niterations < 1 || !(c->R0 > 0) || !(c->approx_frequency_max > 0) ||
        !c->sf_filename[0])
        return fail(HH_E_INVALID, "MaxTrajectoryLength, sampling_time, cvode_tolerance, total_trajectories, niterations, "
                                  "R0, ApproximateFrequencyMax and sf_filename must be set");
    if (c->average_time_between_collisions > 0)
        return fail(HH_E_UNSUPPORTED, "the Poisson free-path cut-off (average_time_between_collisions) needs a "
                                      "requantised rotor, which is not part of this engine");
    if (c->initialM0_npoints < 2 || c->initialM2_npoints < 2)
        return fail(HH_E_INVALID, "PR_MU runs the preliminary M0 and M2: initialM0_npoints and initialM2_npoints "
                                  "must be set");
    if (c->pesmin == 0 || !(c->sampler_Rmin > 0 && c->sampler_Rmax > c->sampler_Rmin))
        return fail(HH_E_INVALID, "pesmin, sampler_Rmin and sampler_Rmax must be set");
    hh_system sys{};
    int rc = resolve_system(e, c, hs, &sys);
    if (rc < 0) return rc;
    const long L = c->max_trajectory_length;
    const double dt = c->sampling_time, T = c->temperature;
    const double ATU = 2.4188843265857e-17, LightSpeed = 2.99792458e8, LightSpeed_cm = 2.99792458e10;   // src/constants.h
    const double frequency_step = 1.0 / (dt * ATU) / LightSpeed_cm / (double)L;
    const double theoretical_frequency_max = 1.0 / (LightSpeed * 100.0 * dt * ATU) / 2.0;
    if (c->approx_frequency_max > theoretical_frequency_max)
        return fail(HH_E_INVALID, "Requested maximum frequency (%.3e) should be less than 1/2 of the maximum signal "
                                  "frequency, which is (%.3e)", c->approx_frequency_max, theoretical_frequency_max);
    const int nfreq = (int)(c->approx_frequency_max / frequency_step) + 1;
    const double max_frequency = frequency_step * (nfreq - 1);
    Series sf;
    sf.sf = true;
    sf.T = T;
    sf.x.resize(nfreq);
    sf.y.assign(nfreq, 0.0);
    const double nu_step = nfreq > 1 ? (max_frequency - 0.0) / (nfreq - 1) : 0.0;   // linspace (src/hawaii.c:3994-4012)
    for (int k = 0; k < nfreq; ++k) sf.x[k] = 0.0 + nu_step * k;
    if (!o.quiet) {
        std::printf("Calculating spectral function using pr/mu representation at T = %.2f\n", T);
        std::printf("Theoretical maximum frequency with provided parameters: %.3e cm-1\n", theoretical_frequency_max);
        std::printf("The frequency step with provided parameters:            %.3e cm-1\n", frequency_step);
        std::printf("The spectral function will be calculated in the range [0 .. %.3e cm-1] for %d points\n",
                    max_frequency, nfreq);
    }
    const int ps = c->pair_state ? c->pair_state : HH_PAIR_FREE;
    double ppf = c->ppf;
    if (ppf <= 0) {
        // src/hawaii.c:3424-3444: 15 iterations of 2e6 points
        const double pf_an = hh_pf_analytic(&sys, T);
        if (pf_an < 0) return (int)HH_E_INVALID;
        hh_vegas_params vp{HH_INTEGRAND_PF, ps, T, c->sampler_Rmin, c->sampler_Rmax, 15, 2000000L, o.seed ? o.seed : 1u};
        hh_vegas_result vr{};
        rc = hh_vegas(e, &sys, &vp, &vr);
        if (rc < 0) return rc;
        ppf = vr.value / pf_an;
        if (!o.quiet) {
            std::printf("Full partition function Q/V = %.12lf\n", pf_an);
            std::printf("Partial partition function Qp/V = %.12lf\n", vr.value);
            std::printf("Ratio of partial partition to full partition function: %.5e\n\n", ppf);
        }
    }
    const uint32_t seed = o.seed ? o.seed : good_seed();
    if (!o.quiet) std::printf("Generator seed is set to %u\n", seed);
    const int qp = hh_qp_size(sys.system);
    const long m0n = c->initialM0_npoints, m2n = c->initialM2_npoints;
    const long n_iter = c->total_trajectories / c->niterations;
    if (n_iter <= 0) return fail(HH_E_INVALID, "total_trajectories < niterations");
    // preliminary moments from the rejection sampler's points
    hh_sampler smp{HH_IC_REJECTION, ps, T, c->sampler_Rmin, c->sampler_Rmax, c->pesmin, 0.0};
    std::vector<double> ic((size_t)std::max(m0n + m2n, n_iter) * qp);
    long draws = 0;
    long lrc = hh_sample_ics_from(&sys, &smp, seed, io ? io->skip_draws : 0, m0n + m2n, hs.pes, ic.data(), &draws);
    if (lrc < 0) return (int)lrc;
    const long nmax = std::max(m0n + m2n, n_iter);
    double *ic_dev = nullptr, *out_dev = nullptr, *sf_dev = nullptr;
    std::vector<double> soa((size_t)nmax * qp);
    auto upload = [&](long first, long n) -> int {
        for (long t = 0; t < n; ++t)
            for (int i = 0; i < qp; ++i) soa[(size_t)i * n + t] = ic[(size_t)(first + t) * qp + i];
        if (hipMemcpy(ic_dev, soa.data(), (size_t)n * qp * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
            return fail(HH_E_DEVICE, "upload of initial conditions failed");
        return (int)HH_OK;
    };
    int code = HH_OK;
    do {
        if (hipMalloc(&ic_dev, (size_t)nmax * qp * sizeof(double)) != hipSuccess ||
            hipMalloc(&out_dev, (size_t)3 * std::max(m0n, m2n) * sizeof(double)) != hipSuccess ||
            hipMalloc(&sf_dev, (size_t)nfreq * sizeof(double)) != hipSuccess ||
            hipMemset(sf_dev, 0, (size_t)nfreq * sizeof(double)) != hipSuccess) {
            code = fail(HH_E_NOMEM, "device allocation failed");
            break;
        }
        double prelim_m0 = 0.0, prelim_m0s = 0.0, prelim_m2 = 0.0, prelim_m2s = 0.0;
        {
            if ((code = upload(0, m0n)) < 0) break;
            if ((code = hh_eval(e, &sys, HH_EVAL_DIPOLE, ic_dev, m0n, out_dev)) < 0) break;
            std::vector<double> d((size_t)3 * m0n), f((size_t)m0n);
            if (hipMemcpy(d.data(), out_dev, d.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
                code = fail(HH_E_DEVICE, "download failed");
                break;
            }
            for (long k = 0; k < m0n; ++k) f[k] = d[k] * d[k] + d[m0n + k] * d[m0n + k] + d[2 * m0n + k] * d[2 * m0n + k];
            welford(f, kZeroCoeff * ppf, &prelim_m0, &prelim_m0s);
            if ((code = upload(m0n, m2n)) < 0) break;
            if ((code = hh_eval(e, &sys, HH_EVAL_DIPOLE_RATE, ic_dev, m2n, out_dev)) < 0) break;
            f.resize(m2n);
            if (hipMemcpy(f.data(), out_dev, f.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
                code = fail(HH_E_DEVICE, "download failed");
                break;
            }
            welford(f, kSecondCoeff * ppf, &prelim_m2, &prelim_m2s);
            if (!o.quiet) {
                std::printf("M0 = %.10e +/- %.10e [%.10e ... %.10e]\n", prelim_m0, prelim_m0s, prelim_m0 - prelim_m0s,
                            prelim_m0 + prelim_m0s);
                std::printf("Error: %.3f%%\n\n", prelim_m0s / prelim_m0 * 100.0);
                std::printf("M2 = %.10e +/- %.10e [%.10e ... %.10e]\n", prelim_m2, prelim_m2s, prelim_m2 - prelim_m2s,
                            prelim_m2 + prelim_m2s);
                std::printf("Error: %.3f%%\n\n", prelim_m2s / prelim_m2 * 100.0);
            }
        }
        hh_prmu_params pp{};
        const bool mixed = o.precision == HH_PREC_MIXED;
        pp.rtol = pp.atol = mixed ? (o.rtol_mixed > 0 ? o.rtol_mixed : 1e-10) : c->cvode_tolerance;
        pp.orbit_rtol = c->cvode_tolerance;   // orbiting pairs: FP64 at the configured tolerance
        pp.dt = dt;
        pp.R0 = c->R0;
        pp.L = (int)L;
        pp.nfreq = nfreq;
        pp.precision = mixed ? HH_PREC_MIXED : HH_PREC_FP64;
        pp.max_steps = 50000;
        hh_sampler inj{HH_IC_R0_INJECT, ps, T, c->sampler_Rmin, c->sampler_Rmax, c->pesmin, c->R0};
        std::vector<int> counted((size_t)n_iter);
        double ntraj = 0.0;
        for (long iter = 0; iter < c->niterations && code == HH_OK; ++iter) {
            long need = n_iter;
            int idle = 0;
            while (need > 0 && code == HH_OK) {
                lrc = hh_sample_ics_from(&sys, &inj, seed, draws, need, hs.pes, ic.data(), &draws);
                if (lrc < 0) {
                    code = (int)lrc;
                    break;
                }
                if ((code = upload(0, need)) < 0) break;
                long nc = 0;
                hh_cf_stats st;
                if ((code = hh_prmu(e, &sys, &pp, ic_dev, need, sf_dev, counted.data(), &nc, &st)) < 0) break;
                if (nc == 0 && ++idle >= 100) {
                    code = fail(HH_E_INVALID, "no trajectory ended beyond R0 within %ld samples in 100 draws; "
                                              "increase MaxTrajectoryLength", L);
                    break;
                }
                if (nc < need && !o.quiet)
                    std::printf("INFO: %ld trajectories exceed maximum length; drawing new ones in their place\n",
                                need - nc);
                need -= nc;
            }
            if (code != HH_OK) break;
            ntraj += (double)n_iter;
            if (hipMemcpy(sf.y.data(), sf_dev, nfreq * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
                code = fail(HH_E_DEVICE, "download failed");
                break;
            }
            double m0_est = 0.0, m2_est = 0.0;
            (void)hh_sf_moment(sf.x.data(), sf.y.data(), nfreq, ntraj, 0, &m0_est);
            (void)hh_sf_moment(sf.x.data(), sf.y.data(), nfreq, ntraj, 2, &m2_est);
            if (!o.quiet) {
                std::printf("ITERATION %ld/%ld: accumulated %d trajectories. Saving temporary result to '%s'\n", iter + 1,
                            c->niterations, (int)ntraj, c->sf_filename);
                std::printf("M0 ESTIMATE FROM SF: %.5e, PRELIMINARY M0 ESTIMATE: %.5e, diff: %.3f%%\n", m0_est,
                            prelim_m0, (m0_est - prelim_m0) / prelim_m0 * 100.0);
                std::printf("M2 ESTIMATE FROM SF: %.5e, PRELIMINARY M2 ESTIMATE: %.5e, diff: %.3f%%\n\n", m2_est,
                            prelim_m2, (m2_est - prelim_m2) / prelim_m2 * 100.0);
                std::fflush(stdout);
            }
            // written normalised by the trajectory count (write_spectral_function_ext, src/hawaii.c:4230-4265)
            Series out = sf;
            out.ntraj = ntraj;
            for (int k = 0; k < nfreq; ++k) out.y[k] = sf.y[k] / ntraj;
            code = write_series(c->sf_filename, out);
            if (code == HH_OK && io && io->sf) *io->sf = out;
        }
        if (io) io->draws_used = draws;
    } while (false);
    if (ic_dev) (void)hipFree(ic_dev);
    if (out_dev) (void)hipFree(out_dev);
    if (sf_dev) (void)hipFree(sf_dev);
    return code;
}

}  // namespace conf
}  // namespace hh

extern "C" {

int hh_conf_parse(const char* path, hh_conf* out)
{
    if (!path || !out) return hh::fail(HH_E_INVALID, "hh_conf_parse: null argument");
    std::ifstream in(path);
    if (!in) return hh::fail(HH_E_INVALID, "could not open '%s'", path);
    std::stringstream ss;
    ss << in.rdbuf();
    return hh::conf::parse_text(ss.str(), path, out);
}

int hh_conf_run(const hh_conf* conf, const hh_conf_run_opts* opts)
{
    using namespace hh::conf;
    if (!conf) return hh::fail(HH_E_INVALID, "hh_conf_run: null configuration");
    hh_conf_run_opts o{};
    if (opts) o = *opts;
    HostSurfaces hs;
    int rc = HH_OK;
    if (conf->calculation_type != HH_CALC_PROCESSING && (rc = load_surfaces(conf, hs)) < 0) return rc;
    hh_engine* e = nullptr;
    rc = hh_engine_create(o.device, &e);
    if (rc < 0) return rc;
    if (conf->calculation_type == HH_CALC_CORRELATION_SINGLE) {
        rc = run_correlation_single(e, conf, o, hs, nullptr);
        if (rc == HH_OK) rc = run_processing(e, conf, o.quiet != 0);
    } else if (conf->calculation_type == HH_CALC_PROCESSING) {
        rc = run_processing(e, conf, o.quiet != 0);
    } else if (conf->calculation_type == HH_CALC_PHASE_SPACE_M0) {
        rc = run_phase_space_m0(e, conf, o, hs);
    } else if (conf->calculation_type == HH_CALC_PHASE_SPACE_M2) {
        rc = run_phase_space_m2(e, conf, o, hs);
    } else if (conf->calculation_type == HH_CALC_PR_MU) {
        rc = run_prmu(e, conf, o, hs, nullptr);
    } else if (conf->calculation_type == HH_CALC_CORRELATION_ARRAY) {
        rc = run_correlation_array(e, conf, o, hs, nullptr);
    } else {
        rc = hh::fail(HH_E_INVALID, "CALCULATION_TYPE is not set");
    }
    hh_engine_destroy(e);
    return rc;
}

}  // extern "C"
