// ORACLE — test infrastructure only (see orc.hpp).
// Grid generator and body front end, restating
//   src/io/parseDomainFile.cu:28-222 (stretched-grid rule, with its int truncations)
//   src/io/parseDomainFile.cu:355-378 (xu, yv)
//   src/io/parseBodiesFile.cu:145-178 (circle points)
#include <cmath>
#include "orc.hpp"

namespace orc {

// cells in one subdomain, first pass of parseDomainFile.cu:41-66 (int() truncation)
static int block_cells(double start, double end, double mid_h, double r) {
    int numCells;
    double a;
    if (r < 1) {
        a = mid_h / ((start - end) * (r - 1));
        numCells = (int)(std::log(1 / a + 1) / std::log(r));
    } else if (r == 1) {
        numCells = (int)((end - start) / mid_h);
    } else {
        a = mid_h / ((end - start) * (r - 1));
        numCells = (int)(std::log(1 / a + 1) / std::log(r));
    }
    return numCells;
}

int axis_cells(double start, double mid_h, int nsub, const double *ends, const double *ratios) {
    int count = 0;
    for (int i = 0; i < nsub; i++) {
        count += block_cells(start, ends[i], mid_h, ratios[i]);
        start = ends[i];
    }
    return count;
}

// second pass, parseDomainFile.cu:80-222
void axis_fill(double start, double mid_h, int nsub, const double *ends, const double *ratios,
               double *x, double *dx) {
    int beg = 0;
    for (int i = 0; i < nsub; i++) {
        const double end = ends[i], r = ratios[i];
        const int n = block_cells(start, end, mid_h, r);
        if (std::fabs(r - 1.0) < 1.0e-6) {  // uniform block
            const double h = (end - start) / n;
            for (int j = beg; j < beg + n; j++) {
                dx[j] = h;
                x[j] = (j == beg) ? start + h / 2 : x[j - 1] + 0.5 * dx[j] + 0.5 * dx[j - 1];
            }
        } else if (end < 0) {  // block left of / below the core: built backwards from `end`
            const double h = (end - start) * (r - 1) / (std::pow(r, n) - 1);
            int jdx = beg;
            for (int j = beg + n - 1; j >= beg; j--) {
                if (j == beg + n - 1) {
                    x[j] = end - h / 2.0;
                    dx[j] = h;
                } else {
                    dx[j] = h * std::pow(r, jdx - beg);
                    x[j] = x[j + 1] - 0.5 * dx[j] - 0.5 * dx[j + 1];
                }
                jdx++;
            }
        } else {  // growing block
            const double h = (end - start) * (r - 1) / (std::pow(r, n) - 1);
            for (int j = beg; j < beg + n; j++) {
                if (j == beg) {
                    x[j] = start + h * std::pow(r, j - beg) / 2;
                    dx[j] = h * std::pow(r, j - beg);
                } else {
                    dx[j] = h * std::pow(r, j - beg);
                    x[j] = x[j - 1] + 0.5 * dx[j] + 0.5 * dx[j - 1];
                }
            }
        }
        beg += n;
        start = end;
    }
}

}  // namespace orc

extern "C" {

int orc_axis_cells(double start, double mid_h, int nsub, const double *ends, const double *ratios) {
    return orc::axis_cells(start, mid_h, nsub, ends, ratios);
}

void orc_axis_fill(double start, double mid_h, int nsub, const double *ends, const double *ratios,
                   double *x, double *dx) {
    orc::axis_fill(start, mid_h, nsub, ends, ratios, x, dx);
}

// circle body, parseBodiesFile.cu:145-178; returns the point count
int orc_circle_points(double cx, double cy, double R, int numPoints, int selfSize, double mid_h,
                      double *X, double *Y) {
    if (selfSize) {
        numPoints = (int)(2.0 * R * M_PI / mid_h);
        if (numPoints % 2 != 0) numPoints++;
    }
    // separate loops keep the compiler from fusing cos and sin into one sincos call,
    // whose last bit can differ from cos alone
    if (X && Y) {
        for (int i = 0; i < numPoints; i++) X[i] = cx + R * std::cos(i * 2 * M_PI / numPoints);
        for (int i = 0; i < numPoints; i++) Y[i] = cy + R * std::sin(i * 2 * M_PI / numPoints);
    }
    return numPoints;
}
}
