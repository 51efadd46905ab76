// Configuration results file of the last converged configuration, in the text
// layout of PostFiles::WriteConfigurationResults (src/PostFiles.cpp:1429-1464):
//   "RESULTS - NODES", a blank line, one Node::WriteResults line per node
//   (src/Node.cpp:208-219: number, X Y Z of the converged coordinates, RX RY RZ
//   of the Euler rotation vector theta = 2 atan(|alpha|/2) alpha/|alpha|,
//   src/Node.cpp:352-360), then "RESULTS - ELEMENTS" and one line per element in
//   element-number order: Beam_1::WriteResults (src/Beam_1.cpp:444-458) and
//   Shell_1::WriteResults (src/Shell_1.cpp:624-666).
//
// The element lines carry Gauss-point means of the back-rotated stresses and
// strains, evaluated on the device from the committed Gauss-point state, and
// the element's reference and current positions.  "Coor" adds the nodes'
// incremental displacements to their converged coordinates as the reference
// does; after a commit those still hold the last step's increment (the
// reference clears them at the next step's start, src/Solution.cpp:798-814),
// so the reference's "Coor" counts that increment twice, and so does this file.
#include "gf_internal.h"

#include <algorithm>
#include <cstdio>
#include <vector>

extern "C" int gf_read_element_results(gf_model *m, double *beam_out, double *shell_out)
{
    if (!m) return GF_E_ARG;
    const size_t nb = (size_t)m->nb, ns = (size_t)m->ns;
    const size_t nbv = 12 * nb, nsv = 24 * ns;
    if ((beam_out ? nbv : 0) + (shell_out ? nsv : 0) == 0) return GF_OK;
    double *d_out = nullptr;
    GF_TRY(gf_pool_alloc((void **)&d_out, sizeof(double) * (nbv + nsv)));
    std::vector<double> h(nbv + nsv);
    int st = GF_OK;
    if (beam_out) st = gf_beam_results(m, d_out);
    if (st == GF_OK && shell_out) st = gf_shell_results(m, d_out + nbv);
    if (st == GF_OK && hipMemcpyAsync(h.data(), d_out, sizeof(double) * h.size(), hipMemcpyDeviceToHost, m->stream) !=
                           hipSuccess) {
        gf_set_error("element results: copy to host failed");
        st = GF_E_DEVICE;
    }
    if (hipStreamSynchronize(m->stream) != hipSuccess && st == GF_OK) {
        gf_set_error("element results: stream synchronisation failed");
        st = GF_E_DEVICE;
    }
    gf_pool_free(d_out);
    if (st != GF_OK) return st;
    // colour-sorted position -> description order
    if (beam_out)
        for (size_t s = 0; s < nb; s++)
            std::copy(&h[12 * s], &h[12 * s] + 12, beam_out + 12 * (size_t)m->beam_order[s]);
    if (shell_out)
        for (size_t s = 0; s < ns; s++)
            std::copy(&h[nbv + 24 * s], &h[nbv + 24 * s] + 24, shell_out + 24 * (size_t)m->shell_order[s]);
    return GF_OK;
}

extern "C" int gf_write_results(gf_model *m, const char *path)
{
    if (!m || !path) return GF_E_ARG;
    const int nn = m->nn, nb = m->nb, ns = m->ns;
    std::vector<double> nodes(6 * (size_t)nn), disp(6 * (size_t)nn);
    std::vector<double> bres(12 * (size_t)nb), sres(24 * (size_t)ns);
    GF_TRY(gf_read_nodes(m, nodes.data()));
    GF_TRY(gf_read_disp(m, disp.data()));
    GF_TRY(gf_read_element_results(m, bres.data(), sres.data()));
    // element number -> (type, index)
    std::vector<int> by_number((size_t)nb + ns + 1, -1);
    for (int e = 0; e < nb; e++) by_number[m->beam_number[e]] = e;
    for (int e = 0; e < ns; e++) by_number[m->shell_number[e]] = nb + e;

    FILE *f = fopen(path, "w");
    if (!f) {
        gf_set_error(std::string("cannot open ") + path);
        return GF_E_ARG;
    }
    fprintf(f, "RESULTS - NODES\n\n");
    for (int i = 0; i < nn; i++) {
        const double *v = &nodes[6 * (size_t)i];
        fprintf(f, "Node\t%d\tX\t%.6e\tY\t%.6e\tZ\t%.6e\tRX\t%.6e\tRY\t%.6e\tRZ\t%.6e\n", i + 1, v[0], v[1], v[2], v[3],
                v[4], v[5]);
    }
    fprintf(f, "\n\nRESULTS - ELEMENTS\n\n");
    const double *xr = m->h_nodes_ref.data();
    for (int num = 1; num <= nb + ns; num++) {
        const int e = by_number[num];
        if (e < nb) {
            const double *r = &bres[12 * (size_t)e];
            const int c = m->h_beam_conn[3 * (size_t)e + 1];  // node 2 of Beam_1
            const double *x = &nodes[6 * (size_t)c], *u = &disp[6 * (size_t)c];
            // Pipe_1::WriteResults (src/Pipe_1.cpp:586-599) prints the same columns
            fprintf(f,
                    "%s\t%d\tStress_r\t%.6e\t%.6e\t%.6e\t%.6e\t%.6e\t%.6e\tStrain_r\t%.6e\t%.6e\t%.6e\t%.6e\t%.6e\t%.6e"
                    "\tRef_Coor\t%.6e\t%.6e\t%.6e\tCoor\t%.6e\t%.6e\t%.6e\n",
                    m->beam_kind[e] == GF_BEAM_KIND_PIPE1 ? "Pipe_1" : "Beam_1", num, r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7], r[8], r[9], r[10], r[11], xr[3 * c],
                    xr[3 * c + 1], xr[3 * c + 2], x[0] + u[0], x[1] + u[1], x[2] + u[2]);
        } else {
            const int s = e - nb;
            const double *r = &sres[24 * (size_t)s];
            const int *cn = &m->h_shell_conn[6 * (size_t)s];  // corners: nodes 1-3 of Shell_1
            double ref[3], cur[3];
            for (int k = 0; k < 3; k++) {
                ref[k] = (xr[3 * cn[0] + k] + xr[3 * cn[1] + k] + xr[3 * cn[2] + k]) / 3;
                cur[k] = (nodes[6 * (size_t)cn[0] + k] + nodes[6 * (size_t)cn[1] + k] + nodes[6 * (size_t)cn[2] + k]) / 3 +
                         (disp[6 * (size_t)cn[0] + k] + disp[6 * (size_t)cn[1] + k] + disp[6 * (size_t)cn[2] + k]) / 3;
            }
            fprintf(f,
                    "Shell_1\t%d\tn1\t%.6e\t%.6e\t%.6e\tm1\t%.6e\t%.6e\t%.6e\tn2\t%.6e\t%.6e\t%.6e\tm2\t%.6e\t%.6e\t%.6e"
                    "\teta1\t%.6e\t%.6e\t%.6e\tkappa1\t%.6e\t%.6e\t%.6e\teta2\t%.6e\t%.6e\t%.6e\tkappa2\t%.6e\t%.6e\t%.6e"
                    "\tRef_Coor\t%.6e\t%.6e\t%.6e\tCoor\t%.6e\t%.6e\t%.6e\n",
                    num, r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7], r[8], r[9], r[10], r[11], r[12], r[13], r[14],
                    r[15], r[16], r[17], r[18], r[19], r[20], r[21], r[22], r[23], ref[0], ref[1], ref[2], cur[0],
                    cur[1], cur[2]);
        }
    }
    const bool ok = fclose(f) == 0;
    if (!ok) {
        gf_set_error(std::string("write failed: ") + path);
        return GF_E_ARG;
    }
    return GF_OK;
}
