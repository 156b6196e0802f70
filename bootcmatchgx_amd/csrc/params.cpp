// params.cpp -- settings: defaults, .properties parsing and enum strings.
//
// Mirrors struct params (src/config/Params.h:238-291) and
// Params::initFromPropertiesFile / parse_properties_file
// (src/config/Params.cpp:272-404) including their quirks: the key/value split
// is at the last '=' before any '#', the value keeps everything after that
// '=' up to the end of the line (a trailing comment included) before
// whitespace trimming, and unknown keys are ignored.  Errors are status codes
// instead of exit().
#include <cctype>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/bcmgx.h"

namespace {

struct EnumTab {
    const char* kind;
    const char* const* names;
    int count;  // entries before INVALID; INVALID == count
};
const char* const SOLVERS[] = {"CGHS", "FCG", "CGS", "PIPELINED_CGS", "CGS_CUBLAS"};
const char* const BOOTS[] = {"MULTIPLICATIVE", "SYMMETRIZED_MULTIPLICATIVE", "ADDITIVE"};
const char* const MATCHES[] = {"SUITOR"};
const char* const CYCLES[] = {"V_CYCLE", "H_CYCLE", "W_CYCLE", "VARIABLE_V_CYCLE"};
const char* const COARSES[] = {"L1_JACOBI"};
const char* const RELAXES[] = {"L1_JACOBI"};
const char* const PRECS[] = {"NONE", "L1_JACOBI", "BCMG", "AFSAI"};
const char* const PRECISIONS[] = {"FP64", "FP32"};
const EnumTab TABS[] = {
    {"solver", SOLVERS, 5},     {"bootstrap_type", BOOTS, 3}, {"matchtype", MATCHES, 1},
    {"cycle_type", CYCLES, 4},  {"coarse_solver", COARSES, 1}, {"relax_type", RELAXES, 1},
    {"preconditioner", PRECS, 4}, {"precision", PRECISIONS, 2},
};

const EnumTab* find_tab(const char* kind) {
    for (const auto& t : TABS)
        if (std::strcmp(t.kind, kind) == 0) return &t;
    return nullptr;
}

void trim(std::string& s) {
    size_t b = 0, e = s.size();
    while (b < e && std::isspace((unsigned char)s[b])) b++;
    while (e > b && std::isspace((unsigned char)s[e - 1])) e--;
    s = s.substr(b, e - b);
}

bool to_int(const std::string& v, int* out) {
    try {
        size_t pos = 0;
        *out = std::stoi(v, &pos);
        return true;
    } catch (...) {
        return false;
    }
}
bool to_double(const std::string& v, double* out) {
    try {
        size_t pos = 0;
        *out = std::stod(v, &pos);
        return true;
    } catch (...) {
        return false;
    }
}

}  // namespace

extern "C" {

int bcmgx_params_default(bcmgx_params* p) {
    if (!p) return BCMGX_ERR_ARG;
    p->mem_alloc_size = 2000000;  // GLOB_MEM_ALLOC_SIZE
    p->error = 0;
    p->solver = BCMGX_SOLVER_CGHS;
    p->itnlim = 2000;
    p->rtol = 1.e-6;
    p->dispnorm = 1;
    p->preconditioner = BCMGX_PREC_NONE;
    p->bootstrap_type = BCMGX_BOOT_MULTIPLICATIVE;
    p->max_hrc = 1;
    p->conv_ratio = 0.8;
    p->matchtype = BCMGX_MATCH_SUITOR;
    p->aggrsweeps = 2;
    p->aggr_type = 0;
    p->max_levels = 39;
    p->cycle_type = BCMGX_CYCLE_V;
    p->coarse_solver = BCMGX_COARSE_L1_JACOBI;
    p->relax_type = BCMGX_RELAX_L1_JACOBI;
    p->relaxnumber_coarse = 20;
    p->prerelax_sweeps = 4;
    p->postrelax_sweeps = 4;
    p->l1jacsweeps = 4;
    p->sstep = 1;
    p->stop_criterion = 1;
    p->ru_res = 1;
    p->rec_res_int = 0;
    p->max_coarse_size = 0;
    p->precision = BCMGX_PRECISION_FP64;
    p->use_graph = 1;
    p->autotune = 1;
    p->check_every = 8;
    p->fcg_z32 = 1;
    p->post_galerkin = 1;
    return BCMGX_OK;
}

int bcmgx_enum_from_string(const char* kind, const char* str) {
    const EnumTab* t = kind ? find_tab(kind) : nullptr;
    if (!t || !str) return BCMGX_ERR_ARG;
    for (int i = 0; i < t->count; i++)
        if (std::strcmp(t->names[i], str) == 0) return i;
    return t->count;  // INVALID
}

const char* bcmgx_enum_to_string(const char* kind, int value) {
    const EnumTab* t = kind ? find_tab(kind) : nullptr;
    if (!t) return nullptr;
    if (value >= 0 && value < t->count) return t->names[value];
    return "INVALID";
}

int bcmgx_params_set(bcmgx_params* p, const char* key_c, const char* value_c) {
    if (!p || !key_c || !value_c) return BCMGX_ERR_ARG;
    std::string key(key_c), v(value_c);
    trim(key);
    trim(v);
    int iv = 0;
    double dv = 0;
    auto seti = [&](int* f) -> int { return to_int(v, f) ? BCMGX_OK : BCMGX_ERR_PARAM; };
    auto setd = [&](double* f) -> int { return to_double(v, f) ? BCMGX_OK : BCMGX_ERR_PARAM; };
    (void)iv;
    (void)dv;
    if (key == "rhsfile" || key == "solfile") return BCMGX_OK;  // strings the solve path does not use
    if (key == "solver") { p->solver = bcmgx_enum_from_string("solver", v.c_str()); return BCMGX_OK; }
    if (key == "itnlim") return seti(&p->itnlim);
    if (key == "rtol") return setd(&p->rtol);
    if (key == "dispnorm") return seti(&p->dispnorm);
    if (key == "preconditioner") { p->preconditioner = bcmgx_enum_from_string("preconditioner", v.c_str()); return BCMGX_OK; }
    if (key == "bootstrap_type") { p->bootstrap_type = bcmgx_enum_from_string("bootstrap_type", v.c_str()); return BCMGX_OK; }
    if (key == "max_hrc") return seti(&p->max_hrc);
    if (key == "conv_ratio") return setd(&p->conv_ratio);
    if (key == "matchtype") { p->matchtype = bcmgx_enum_from_string("matchtype", v.c_str()); return BCMGX_OK; }
    if (key == "aggrsweeps") return seti(&p->aggrsweeps);
    if (key == "aggr_type") return seti(&p->aggr_type);
    if (key == "max_levels") return seti(&p->max_levels);
    if (key == "cycle_type") { p->cycle_type = bcmgx_enum_from_string("cycle_type", v.c_str()); return BCMGX_OK; }
    if (key == "coarse_solver") { p->coarse_solver = bcmgx_enum_from_string("coarse_solver", v.c_str()); return BCMGX_OK; }
    if (key == "relax_type") { p->relax_type = bcmgx_enum_from_string("relax_type", v.c_str()); return BCMGX_OK; }
    if (key == "relaxnumber_coarse") return seti(&p->relaxnumber_coarse);
    if (key == "prerelax_sweeps") return seti(&p->prerelax_sweeps);
    if (key == "postrelax_sweeps") return seti(&p->postrelax_sweeps);
    if (key == "l1jacsweeps") return seti(&p->l1jacsweeps);
    if (key == "sstep") return seti(&p->sstep);
    if (key == "stop_criterion") return seti(&p->stop_criterion);
    if (key == "ru_res") return seti(&p->ru_res);
    if (key == "rec_res_int") return seti(&p->rec_res_int);
    // extension keys
    if (key == "max_coarse_size") return seti(&p->max_coarse_size);
    if (key == "precision") {
        int e = bcmgx_enum_from_string("precision", v.c_str());
        if (e < 0 || e > 1) return BCMGX_ERR_PARAM;
        p->precision = e;
        return BCMGX_OK;
    }
    if (key == "use_graph") return seti(&p->use_graph);
    if (key == "autotune") return seti(&p->autotune);
    if (key == "check_every") return seti(&p->check_every);
    if (key == "fcg_z32") return seti(&p->fcg_z32);
    if (key == "post_galerkin") return seti(&p->post_galerkin);
    return BCMGX_OK;  // unknown keys are ignored (Params.cu:345-398)
}

int bcmgx_params_from_properties(const char* path, bcmgx_params* p) {
    if (!path || !p) return BCMGX_ERR_ARG;
    bcmgx_params_default(p);
    FILE* fp = std::fopen(path, "r");
    if (!fp) return BCMGX_ERR_IO;
    char buf[1025];
    int status = BCMGX_OK;
    while (std::fgets(buf, 1024, fp)) {
        const char* hash = std::strchr(buf, '#');
        int len = hash ? (int)(hash - buf) : (int)std::strlen(buf);
        if (len > 0 && buf[len - 1] == '\n') len--;
        bool empty = true;
        int eq = -1;
        for (int i = 0; i < len; i++) {
            if (buf[i] != ' ') empty = false;
            if (buf[i] == '=') eq = i;
        }
        if (empty) continue;
        if (eq < 0) {  // "could not find '='" (Params.cu:306-309)
            status = BCMGX_ERR_IO;
            break;
        }
        std::string key(buf, eq), value(buf + eq + 1);
        int s = bcmgx_params_set(p, key.c_str(), value.c_str());
        if (s != BCMGX_OK) {
            status = s;
            break;
        }
    }
    std::fclose(fp);
    if (status != BCMGX_OK) p->error = 1;
    return status;
}

}  // extern "C"
