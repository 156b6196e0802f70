// bcmgx_solve -- command-line solver over the C ABI (include/bcmgx.h).
//
// The counterpart of the reference's driverSolve (src/example/driverSolve.cu:59-406)
// with the same options and output files: read (-m) or generate (-a, -g/-l)
// the matrix, read the .properties settings (-s), set up the preconditioner,
// solve A x = 1 from x = 0 (driverSolve.cu:241-249), write the info file
// (-i, SolverOut::dump) and the solution (-o, one "%g" value per line as
// Vector::print, vector.cu:378-410).  One process holds the whole matrix: a
// -l grid of P x Q x R blocks is generated in the block order the reference's
// ranks own (laplacian.cu:14-25), so -o writes the vector its aggregate_vector
// would gather.  Options of the reference's profiling and multi-process
// outputs (-p, -P, -M, -x, -B, -O, -S) are accepted and produce nothing.
#include <getopt.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/bcmgx.h"

static const char* USAGE =
    "Usage: %s [--matrix <FILE_NAME> | --laplacian <SIZE> | --laplacian-3d <FILE_NAME>] --settings <FILE_NAME>\n\n"
    "\tYou can specify only one out of the three available options: --matrix, --laplacian-3d and --laplacian\n\n"
    "\t-a, --laplacian <SIZE>                      Generate a 7-point matrix whose size is <SIZE>^3.\n"
    "\t-d, --dump-matrix <FILE_NAME>               Write the input matrix to <FILE_NAME>0.mtx.\n"
    "\t-e, --errlog <FILE_NAME>                    Write the log to <FILE_NAME>0.\n"
    "\t-g, --laplacian-3d-generator [ 7p | 27p ]   Choose laplacian 3d generator (7 points or 27 points).\n"
    "\t-h, --help                                  Print this message.\n"
    "\t-i, --info <FILE_NAME>                      Write info to <FILE_NAME>.\n"
    "\t-l, --laplacian-3d <FILE_NAME>              Read generation parameters (nx ny nz P Q R) from <FILE_NAME>.\n"
    "\t-m, --matrix <FILE_NAME>                    Read the matrix from the Matrix Market file <FILE_NAME>.\n"
    "\t-o, --out <FILE_NAME>                       Write solution to <FILE_NAME>.\n"
    "\t-s, --settings <FILE_NAME>                  Read settings from the .properties file <FILE_NAME>.\n"
    "\t-p, -P, -M, -x, -B, -O, -S                  Profiling and output-directory options: accepted, no output.\n\n";

static FILE* g_log = nullptr;

static void say(const char* fmt, const char* a = nullptr) {
    std::printf(fmt, a);
    std::fflush(stdout);
    if (g_log) {
        std::fprintf(g_log, fmt, a);
        std::fflush(g_log);
    }
}

static int fail(bcmgx_context ctx, int st, const char* what) {
    std::fprintf(stderr, "%s failed: %s (%s)\n", what, bcmgx_error_string(st), bcmgx_last_error(ctx));
    if (g_log) std::fprintf(g_log, "%s failed: %s (%s)\n", what, bcmgx_error_string(st), bcmgx_last_error(ctx));
    return EXIT_FAILURE;
}

// read_laplacian_file (laplacian.cu:974-1020): "name = value" lines, '#' comments
static bool read_lap3d_cfg(const char* path, int prm[6]) {
    static const char* names[6] = {"nx", "ny", "nz", "P", "Q", "R"};
    FILE* fp = std::fopen(path, "r");
    if (!fp) {
        std::fprintf(stderr, "[ERROR] - Laplacian file not found! (%s)\n", path);
        return false;
    }
    for (int i = 0; i < 6; i++) prm[i] = 0;
    char line[1024];
    bool ok = true;
    while (std::fgets(line, sizeof line, fp)) {
        if (line[0] == '#' || line[0] == '\n' || line[0] == '\0') continue;
        char key[64];
        int v;
        if (std::sscanf(line, "%63s = %d", key, &v) != 2) {
            std::fprintf(stderr, "[ERROR] Error reading file %s.\n", path);
            ok = false;
            break;
        }
        for (int i = 0; i < 6; i++)
            if (std::strcmp(key, names[i]) == 0) prm[i] = v;
    }
    std::fclose(fp);
    for (int i = 0; ok && i < 6; i++)
        if (prm[i] < 1) {
            std::fprintf(stderr, "[ERROR] %s: %s missing or not positive\n", path, names[i]);
            ok = false;
        }
    return ok;
}

// CSRMatrixPrintMM: coordinate real general, 1-based
static bool dump_mtx(const std::string& path, int n, int m, int nnz, const int* row, const int* col, const double* val) {
    FILE* fp = std::fopen(path.c_str(), "w");
    if (!fp) return false;
    std::fprintf(fp, "%%%%MatrixMarket matrix coordinate real general\n%d %d %d\n", n, m, nnz);
    for (int i = 0; i < n; i++)
        for (int j = row[i]; j < row[i + 1]; j++) std::fprintf(fp, "%d %d %.17g\n", i + 1, col[j] + 1, val[j]);
    std::fclose(fp);
    return true;
}

int main(int argc, char** argv) {
    enum { NONE, MTX, LAP_3D, LAP } opt = NONE;
    const char* mtx_file = nullptr;
    const char* lap_3d_file = nullptr;
    int points = 27;  // the reference's default generator is LAP_27P
    int lap_n = 0;
    std::string settings, out_file, info_file, log_file, dump_prefix;
    static struct option long_options[] = {
        {"laplacian", required_argument, nullptr, 'a'},   {"out-prefix", required_argument, nullptr, 'B'},
        {"dump-matrix", required_argument, nullptr, 'd'}, {"errlog", required_argument, nullptr, 'e'},
        {"laplacian-3d-generator", required_argument, nullptr, 'g'},
        {"help", no_argument, nullptr, 'h'},              {"info", required_argument, nullptr, 'i'},
        {"laplacian-3d", required_argument, nullptr, 'l'}, {"matrix", required_argument, nullptr, 'm'},
        {"detailed-metrics", required_argument, nullptr, 'M'}, {"out", required_argument, nullptr, 'o'},
        {"out-dir", required_argument, nullptr, 'O'},     {"summary-prof", required_argument, nullptr, 'p'},
        {"detailed-prof", required_argument, nullptr, 'P'}, {"settings", required_argument, nullptr, 's'},
        {"out-suffix", required_argument, nullptr, 'S'},  {"extended-prof", no_argument, nullptr, 'x'},
        {nullptr, 0, nullptr, 0}};
    int ch;
    while ((ch = getopt_long(argc, argv, "a:B:d:e:g:hi:l:m:M:o:O:p:P:s:S:x", long_options, nullptr)) != -1) {
        switch (ch) {
        case 'a': lap_n = std::atoi(optarg); opt = LAP; break;
        case 'd': dump_prefix = optarg; break;
        case 'e': log_file = optarg; break;
        case 'g':
            if (std::strcmp(optarg, "7p") == 0) points = 7;
            else if (std::strcmp(optarg, "27p") == 0) points = 27;
            else points = 0;  // INVALID_GEN
            break;
        case 'i': info_file = optarg; break;
        case 'l': lap_3d_file = optarg; opt = LAP_3D; break;
        case 'm': mtx_file = optarg; opt = MTX; break;
        case 'o': out_file = optarg; break;
        case 's': settings = optarg; break;
        case 'B': case 'M': case 'O': case 'p': case 'P': case 'S': case 'x': break;
        case 'h':
        default: std::printf(USAGE, argv[0]); return EXIT_FAILURE;
        }
    }
    if (opt == NONE || settings.empty() || points == 0 || (opt == LAP && lap_n < 1)) {
        std::printf(USAGE, argv[0]);
        return EXIT_FAILURE;
    }
    if (!log_file.empty()) g_log = std::fopen((log_file + "0").c_str(), "w");

    // ---- settings (Params::initFromPropertiesFile)
    bcmgx_params p;
    bcmgx_params_default(&p);
    if (settings.size() < 11 || settings.compare(settings.size() - 11, 11, ".properties") != 0) {
        std::fprintf(stderr, "%s: only .properties settings files are read\n", settings.c_str());
        return EXIT_FAILURE;
    }
    int st = bcmgx_params_from_properties(settings.c_str(), &p);
    if (st != BCMGX_OK || p.error != 0) return fail(nullptr, st != BCMGX_OK ? st : BCMGX_ERR_PARAM, "reading the settings");

    // ---- matrix
    int n = 0, m = 0, nnz = 0;
    int *row = nullptr, *col = nullptr;
    double* val = nullptr;
    if (opt == MTX) {
        st = bcmgx_read_mtx(mtx_file, &n, &m, &nnz, &row, &col, &val);
        if (st != BCMGX_OK) return fail(nullptr, st, "reading the matrix");
    } else {
        int prm[6] = {lap_n, lap_n, lap_n, 1, 1, 1};
        int pts = 7;  // generateLocalLaplacian3D (laplacian.cu:27-117): the 7-point operator on SIZE^3
        if (opt == LAP_3D) {
            if (!read_lap3d_cfg(lap_3d_file, prm)) return EXIT_FAILURE;
            pts = points;
            std::fprintf(stderr, "Using laplacian 3d %d points generator.\n", pts);
        }
        st = bcmgx_gen_laplacian3d(pts, prm[0], prm[1], prm[2], prm[3], prm[4], prm[5], &n, &nnz, &row, &col, &val);
        if (st != BCMGX_OK) return fail(nullptr, st, "generating the matrix");
        m = n;
    }
    if (!dump_prefix.empty() && !dump_mtx(dump_prefix + "0.mtx", n, m, nnz, row, col, val))
        std::fprintf(stderr, "Error opening %s0.mtx for writing\n", dump_prefix.c_str());

    bcmgx_context ctx = nullptr;
    st = bcmgx_context_create(0, &ctx);
    if (st != BCMGX_OK) return fail(nullptr, st, "creating the context");
    bcmgx_matrix A = nullptr;
    st = bcmgx_csr_create(ctx, n, m, nnz, row, col, val, BCMGX_MEM_HOST, &A);
    bcmgx_free(row);
    bcmgx_free(col);
    bcmgx_free(val);
    if (st != BCMGX_OK) return fail(ctx, st, "copying the matrix to the device");

    // ---- rhs = 1, x0 = 0 (driverSolve.cu:241-249)
    std::vector<double> b((size_t)n, 1.0), x((size_t)n, 0.0);

    say("Setting preconditioner\n");
    bcmgx_prec prec = nullptr;
    st = bcmgx_prec_setup(ctx, A, &p, &prec);
    if (st != BCMGX_OK) return fail(ctx, st, "prec_setup");

    say("Solving\n");
    bcmgx_solver_out out;
    std::memset(&out, 0, sizeof out);
    st = bcmgx_solve(ctx, A, prec, &p, b.data(), x.data(), 1, n, BCMGX_MEM_HOST, &out);
    if (st != BCMGX_OK) return fail(ctx, st, "solve");

    if (!info_file.empty()) {
        say("Dumping output\n");
        st = bcmgx_dump_info(info_file.c_str(), &p, prec, &out);
        if (st != BCMGX_OK) return fail(ctx, st, "writing the info file");
    }
    say("Finalizing preconditioner\n");
    bcmgx_prec_finalize(prec);
    bcmgx_csr_destroy(A);
    bcmgx_context_destroy(ctx);

    if (!out_file.empty()) {
        say("Dumping solution\n");
        FILE* fp = std::fopen(out_file.c_str(), "w");
        if (!fp) {
            std::printf("Error opening %s for writing\n", out_file.c_str());
        } else {
            for (int i = 0; i < n; i++) std::fprintf(fp, "%g\n", x[(size_t)i]);
            std::fclose(fp);
            say("...done.\n");
        }
    }
    say("Shutdown\n");
    if (g_log) std::fclose(g_log);
    return 0;
}
