// bcmgx_compare -- the regression criterion of the reference's test loop
// (compareSolutions, src/example/compareSolutions.cu:20-71): read two vectors
// (one value per line, Vector::load, vector.cu:329-376), print
// ||x1 - x2||_2 / ||x1||_2 and exit 1 when it exceeds the optional threshold.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

static bool load(const char* path, std::vector<double>& v) {
    FILE* fp = std::fopen(path, "r");
    if (!fp) {
        std::fprintf(stdout, "Error opening file %s, errno = %d: %s\n", path, errno, std::strerror(errno));
        return false;
    }
    char line[4096];
    int n = 0;
    while (std::fgets(line, sizeof line, fp)) {
        double x;
        if (std::sscanf(line, "%lf", &x) != 1) {
            std::printf("Error loading vector from %s: Missing value at line %d\n", path, n + 1);
            std::fclose(fp);
            return false;
        }
        v.push_back(x);
        n++;
    }
    std::fclose(fp);
    return true;
}

int main(int argc, char** argv) {
    if (argc < 3 || argc > 4) {
        std::printf("Usage:\n\t%s <FILE_NAME_1.txt> <FILE_NAME_2.txt> [THRESHOLD]\n", argv[0]);
        return EXIT_FAILURE;
    }
    std::vector<double> a, b;
    if (!load(argv[1], a) || !load(argv[2], b)) return 1;
    if (a.size() != b.size()) {
        std::printf("The vectors differ in length (%zu, %zu)\n", a.size(), b.size());
        return 1;
    }
    double den = 0.0, num = 0.0;
    for (size_t i = 0; i < a.size(); i++) {
        den += a[i] * a[i];
        double d = a[i] - b[i];
        num += d * d;
    }
    den = std::sqrt(den);
    if (den == 0.0) {
        std::printf("sol1 must be != <0>\n");
        return 1;
    }
    double result = std::sqrt(num) / den;
    std::printf("Relative error: %lg\n", result);
    if (argc > 3) {
        double threshold = std::atof(argv[3]);
        if (result > threshold) {
            std::printf("Relative error exceeds threshold (%lg)\n", threshold);
            return 1;
        }
    }
    return 0;
}
