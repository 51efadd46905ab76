// Configuration results file of the last converged configuration, in the text
// layout of PostFiles::WriteConfigurationResults (src/PostFiles.cpp:1429-1464):
//   "RESULTS - NODES", a blank line, one Node::WriteResults line per node
//   (src/Node.cpp:208-219: number, X Y Z of the converged coordinates, RX RY RZ
//   of the Euler rotation vector theta = 2 atan(|alpha|/2) alpha/|alpha|,
//   src/Node.cpp:352-360), then "RESULTS - ELEMENTS".  The element lines
//   (Beam_1 / Shell_1 Gauss-point stress averages) are not written: the device
//   path keeps no per-Gauss-point stresses beyond the committed state.
#include "gf_internal.h"

#include <cstdio>
#include <vector>

extern "C" int gf_write_results(gf_model *m, const char *path)
{
    if (!m || !path) return GF_E_ARG;
    std::vector<double> nodes(6 * (size_t)m->nn);
    GF_TRY(gf_read_nodes(m, nodes.data()));
    FILE *f = fopen(path, "w");
    if (!f) {
        gf_set_error(std::string("cannot open ") + path);
        return GF_E_ARG;
    }
    fprintf(f, "RESULTS - NODES\n\n");
    for (int i = 0; i < m->nn; i++) {
        const double *v = &nodes[6 * (size_t)i];
        fprintf(f, "Node\t%d\tX\t%.6e\tY\t%.6e\tZ\t%.6e\tRX\t%.6e\tRY\t%.6e\tRZ\t%.6e\n", i + 1, v[0], v[1], v[2], v[3],
                v[4], v[5]);
    }
    fprintf(f, "\n\nRESULTS - ELEMENTS\n\n");
    const bool ok = fclose(f) == 0;
    if (!ok) {
        gf_set_error(std::string("write failed: ") + path);
        return GF_E_ARG;
    }
    return GF_OK;
}
