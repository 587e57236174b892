// mc_io.hip -- flow-field output in the reference's VTK layout (host code).
//
// FlowSolver<T>::output_data (FlowSolver.inl:18-23) hands the fields to
// VisitWriter<T>::write_flow (include/visit/VisitWriter.hpp:127-234), which
// writes one legacy-ASCII VTK POLYDATA file per rank and output step: the cell
// centres as points, one vertex per point, then the velocity vector and nine
// scalars as point data.  The file is reproduced byte for byte (C++ stream
// formatting: default precision 6, then 11 from the temperature section on,
// as the reference's sticky std::setprecision(11) leaves it), including the
// reference's section labels, under which "flow_ful" holds PRO and
// "flow_pro" holds FUL.
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <string>
#include <vector>

#include "mc_internal.hpp"

namespace {

// field order of the fields argument: U V W P TE ED TEM FUL PRO VARF VARP
enum { VU, VV, VW, VP, VTE, VED, VTEM, VFUL, VPRO, VVARF, VVARP, VN };

void scalar_section(std::ofstream &f, const char *name, int64_t n, const double *v)
{
    f << "SCALARS " << name << " float" << std::endl;
    f << "LOOKUP_TABLE default" << std::endl;
    for (int64_t i = 0; i < n; i++) f << v[i] << "\t";
    f << std::endl;
}

}  // namespace

extern "C" int mc_vtk_write_arrays(const char *path, int64_t n, const double *centres, const double *const *fields)
{
    if (!path || n < 0 || (n > 0 && (!centres || !fields))) return MC_ERR_ARG;
    for (int k = 0; k < VN && n > 0; k++)
        if (!fields[k]) return MC_ERR_ARG;
    std::ofstream f(path);
    if (!f) return MC_ERR_ARG;
    f << "# vtk DataFile Version 3.0 " << std::endl;
    f << "MiniCOMBUST " << std::endl;
    f << "ASCII " << std::endl;
    f << "DATASET POLYDATA " << std::endl;
    f << std::endl << "POINTS " << n << " float" << std::endl;
    for (int64_t i = 0; i < n; i++) {
        f << ((i % 10 == 0) ? "\n" : "  ");
        f << centres[3 * i] << " " << centres[3 * i + 1] << " " << centres[3 * i + 2] << "\t";
    }
    f << std::endl << std::endl;
    f << std::endl << "VERTICES " << n << " " << n * 2 << std::endl;
    for (int64_t i = 0; i < n; i++) f << "1 " << i << "\t";
    f << std::endl;
    f << std::endl << "POINT_DATA " << n << std::endl;
    f << "VECTORS flow_velocity float" << std::endl;
    for (int64_t i = 0; i < n; i++) f << fields[VU][i] << " " << fields[VV][i] << " " << fields[VW][i] << " " << "\t";
    f << std::endl;
    scalar_section(f, "flow_pressure", n, fields[VP]);
    scalar_section(f, "flow_te", n, fields[VTE]);
    scalar_section(f, "flow_ed", n, fields[VED]);
    f << std::setprecision(11);
    scalar_section(f, "flow_tem", n, fields[VTEM]);
    scalar_section(f, "flow_ful", n, fields[VPRO]);   // the reference's labels (VisitWriter.hpp:201-215)
    scalar_section(f, "flow_pro", n, fields[VFUL]);
    scalar_section(f, "flow_varf", n, fields[VVARF]);
    scalar_section(f, "flow_varp", n, fields[VVARP]);
    f.close();
    return f ? MC_OK : MC_ERR_ARG;
}

extern "C" int mc_write_vtk(mc_ctx *x, const char *prefix, int64_t timestep, int rank)
{
    if (!x) return MC_ERR_ARG;
    mc::Ctx *c = &x->c;
    if (!prefix) { c->err = "null prefix"; return MC_ERR_ARG; }
    const int64_t n = c->N;
    MC_HIP(c, hipStreamSynchronize(c->stream));
    std::vector<double> cen(3 * n), t(n);
    const double *src[3] = {c->cx, c->cy, c->cz};
    for (int k = 0; k < 3; k++) {
        MC_HIP(c, hipMemcpy(t.data(), src[k], (size_t)n * 8, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < n; i++) cen[3 * i + k] = t[i];
    }
    static const int ids[VN] = {mc::U, mc::V, mc::W, mc::P, mc::TE, mc::ED, mc::TEM, mc::FUL, mc::PRO, mc::VARF, mc::VARP};
    std::vector<std::vector<double>> v(VN, std::vector<double>(n));
    const double *fp[VN];
    for (int k = 0; k < VN; k++) {
        MC_HIP(c, hipMemcpy(v[k].data(), c->phi[ids[k]], (size_t)n * 8, hipMemcpyDeviceToHost));
        fp[k] = v[k].data();
    }
    // VisitWriter.hpp:131
    const std::string path = std::string(prefix) + "_flow" + std::to_string(rank) + "_timestep" +
                             std::to_string(timestep) + ".vtk";
    const int rc = mc_vtk_write_arrays(path.c_str(), n, cen.data(), fp);
    if (rc != MC_OK) c->err = "cannot write " + path;
    return rc;
}
