// xsload.cpp -- cross-section files beyond read_cross_section (mcc.cu:17-41):
// two-column tables whose numbers may be written m*10^{e} (the notation of
// csf/Ar_conv.txt, whose rows read_cross_section drops), LXCat blocks straight
// into a pic_xs, and read-back of the processes a pic_xs holds. Host code only.
//
// LXCat block layout (the LXCat download format):
//   KEYWORD                      ELASTIC EFFECTIVE ISOTROPIC BACKSCAT EXCITATION
//                                IONIZATION ATTACHMENT
//   target line                  e.g. "Ar", "Ar -> Ar^+"
//   parameter line               mass ratio (elastic kinds) or threshold eV
//                                (excitation, ionization); absent for ATTACHMENT
//   free text                    SPECIES:, PROCESS:, COMMENT:, COLUMNS: ...
//   -----                        a line starting with 5 dashes
//   energy  sigma                rows, eV and m^2
//   -----
// Text outside blocks is ignored.
#include "objects.hpp"
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

using namespace pic;

namespace {

// One number at p: strtod syntax, or mantissa*10^{exponent} / mantissa*10^exponent.
// The second form is rewritten as "<mantissa>e<exponent>" and handed to strtod,
// so both spellings of a value give the same double. Advances p; false if no number.
bool parse_number(const char*& p, double* v) {
    while (*p == ' ' || *p == '\t' || *p == '\r') ++p;
    char* end = nullptr;
    errno = 0;
    const double m = std::strtod(p, &end);
    if (end == p) return false;
    const char* q = end;
    if (q[0] == '*' && q[1] == '1' && q[2] == '0' && q[3] == '^') {
        q += 4;
        const bool brace = *q == '{';
        if (brace) ++q;
        std::string ex;
        if (*q == '+' || *q == '-') ex += *q++;
        while (std::isdigit((unsigned char)*q)) ex += *q++;
        if (ex.empty() || ex == "+" || ex == "-") return false;
        if (brace) {
            if (*q != '}') return false;
            ++q;
        }
        const std::string text = std::string(p, (const char*)end) + "e" + ex;
        char* e2 = nullptr;
        errno = 0;
        *v = std::strtod(text.c_str(), &e2);
        if (e2 != text.c_str() + text.size()) return false;
        p = q;
    } else {
        *v = m;
        p = end;
    }
    return std::isfinite(*v) && errno != ERANGE;
}

// a row "energy sigma" (anything after the two numbers is ignored, as with a
// stream read of two doubles)
bool parse_row(const std::string& line, double* e, double* s) {
    const char* p = line.c_str();
    if (!parse_number(p, e)) return false;
    if (*p != ' ' && *p != '\t') return false;
    return parse_number(p, s);
}

bool is_dashes(const std::string& line) {
    size_t i = 0;
    while (i < line.size() && (line[i] == ' ' || line[i] == '\t')) ++i;
    return line.compare(i, 5, "-----") == 0;
}

std::string first_word(const std::string& line) {
    size_t i = 0;
    while (i < line.size() && std::isspace((unsigned char)line[i])) ++i;
    size_t j = i;
    while (j < line.size() && !std::isspace((unsigned char)line[j])) ++j;
    return line.substr(i, j - i);
}

enum Kind { K_NONE = -1, K_ELASTIC, K_CX, K_EXC, K_ION, K_ATT };

Kind keyword(const std::string& w) {
    if (w == "ELASTIC" || w == "EFFECTIVE" || w == "ISOTROPIC") return K_ELASTIC;
    if (w == "BACKSCAT") return K_CX;
    if (w == "EXCITATION") return K_EXC;
    if (w == "IONIZATION") return K_ION;
    if (w == "ATTACHMENT") return K_ATT;
    return K_NONE;
}

}  // namespace

extern "C" int pic_read_table(const char* path, int cap, double* energies, double* sigmas, int* n) {
    PIC_REQUIRE(path && n, "NULL argument");
    std::ifstream f(path);
    if (!f.is_open()) {
        set_error(std::string("cannot open cross-section file: ") + path);
        return PIC_ERR_IO;
    }
    std::string line;
    int k = 0;
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '#') continue;
        double e, s;
        if (!parse_row(line, &e, &s)) continue;
        if (energies && sigmas) {
            if (k >= cap) {
                set_error("pic_read_table: output capacity too small");
                return PIC_ERR_CAPACITY;
            }
            energies[k] = e;
            sigmas[k] = s;
        }
        ++k;
    }
    *n = k;
    return 0;
}

extern "C" int pic_xs_load_lxcat(pic_xs* xs, const char* path, int species, int product_species, int* n_added,
                                 int* n_skipped) {
    PIC_REQUIRE(xs && path && n_added && n_skipped, "NULL argument");
    PIC_REQUIRE(species >= 0 && species < PIC_MAX_SPECIES, "bad species");
    std::ifstream f(path);
    if (!f.is_open()) {
        set_error(std::string("cannot open LXCat file: ") + path);
        return PIC_ERR_IO;
    }
    std::vector<std::string> lines;
    for (std::string l; std::getline(f, l);) {
        if (!l.empty() && l.back() == '\r') l.pop_back();
        lines.push_back(l);
    }
    std::vector<pic_xs::Proc> add;
    int skipped = 0;
    size_t i = 0;
    const size_t L = lines.size();
    while (i < L) {
        const Kind k = keyword(first_word(lines[i]));
        if (k == K_NONE) {
            ++i;
            continue;
        }
        const size_t head = i;
        auto where = [&](const char* what) {
            set_error(std::string("LXCat ") + path + " line " + std::to_string(head + 1) + ": " + what);
            return PIC_ERR_ARG;
        };
        i += 2;   // keyword and target lines
        double param = 0.0;
        if (k != K_ATT) {
            if (i >= L) return where("block ends before its parameter line");
            const char* p = lines[i].c_str();
            if (!parse_number(p, &param)) return where("no mass ratio or threshold");
            ++i;
        }
        while (i < L && !is_dashes(lines[i])) ++i;
        if (i >= L) return where("no table (missing ----- line)");
        ++i;
        std::vector<double> e, s;
        while (i < L && !is_dashes(lines[i])) {
            double a, b;
            if (!parse_row(lines[i], &a, &b)) return where("bad table row");
            e.push_back(a);
            s.push_back(b);
            ++i;
        }
        if (i >= L) return where("table not closed by a ----- line");
        ++i;
        if (k == K_ATT) {
            ++skipped;   // no attachment process in the collision model
            continue;
        }
        if (e.size() < 2) return where("table shorter than 2 rows");
        for (size_t r = 1; r < e.size(); ++r)
            if (!(e[r] > e[r - 1])) return where("energies must increase");
        pic_xs::Proc pr{species, PIC_PROC_ELASTIC, -1, 0.0, 0.0, e, s};
        if (k == K_CX) pr.type = PIC_PROC_CHARGE_EXCHANGE;
        if (k == K_EXC || k == K_ION) {
            if (!(param > 0.0)) return where("inelastic processes need a positive threshold");
            pr.type = k == K_EXC ? PIC_PROC_EXCITATION : PIC_PROC_IONIZATION;
            pr.thr = param;
        }
        if (k == K_ION) {
            if (product_species < 0 || product_species == species)
                return where("ionization needs a product species different from the projectile");
            pr.product = product_species;
        }
        add.push_back(std::move(pr));
    }
    int have = 0;
    for (const auto& p : xs->procs) have += p.species == species;
    if (have + (int)add.size() > PIC_MAX_PROC) {
        set_error(std::string("LXCat ") + path + ": " + std::to_string(add.size()) + " processes would give species " +
                  std::to_string(species) + " more than 4");
        return PIC_ERR_CAPACITY;
    }
    for (auto& p : add) xs->procs.push_back(std::move(p));
    *n_added = (int)add.size();
    *n_skipped = skipped;
    return 0;
}

extern "C" int pic_xs_process_count(const pic_xs* xs, int* n) {
    PIC_REQUIRE(xs && n, "NULL argument");
    *n = (int)xs->procs.size();
    return 0;
}

extern "C" int pic_xs_get_process(const pic_xs* xs, int k, pic_xs_process* out, int cap, double* energies,
                                  double* sigmas) {
    PIC_REQUIRE(xs && out, "NULL argument");
    PIC_REQUIRE(k >= 0 && k < (int)xs->procs.size(), "process index out of range");
    const auto& p = xs->procs[k];
    out->species = p.species;
    out->type = p.type;
    out->product_species = p.product;
    out->n_points = (int)p.e.size();
    out->threshold_eV = p.thr;
    out->peak_m2 = p.peak;
    if (energies && sigmas) {
        PIC_REQUIRE(cap >= (int)p.e.size(), "output capacity too small");
        for (size_t i = 0; i < p.e.size(); ++i) {
            energies[i] = p.e[i];
            sigmas[i] = p.s[i];
        }
    }
    return 0;
}
