// cf_output.cpp -- the output writer (src/SiloWriter.hpp:61-418 without
// Silo, which is not available): the same data, the owned q and the
// cell-centre velocity interpolated linearly from the faces (computed on
// the device, k_cell_output), and the same naming, a master file per step
// that names the per-rank pieces, written as VTK XML ImageData (.vti) with
// a parallel master (.pvti), which VisIt and ParaView read.
//
//   <dir>/CajitaFluids<step:05d>.pvti              master (DBPutMultimesh/-var)
//   <dir>/raw/CajitaFluidsOutput<group:05d><step:05d>.vti   one piece per rank
//
// A piece holds CellData "quantity" (1 component) and "velocity" (3
// components, the third 0 in 2-D), Float64, raw little-endian appended
// data with UInt64 block headers, and FieldData CYCLE, TIME and DTIME (the
// DBOPT_CYCLE / DBOPT_TIME / DBOPT_DTIME options).
#include <sys/stat.h>

#include <cerrno>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "cf_solver.hpp"

namespace cfmi
{
namespace
{

void make_dir(const std::string& d)
{
    if (d.empty())
        return;
    if (mkdir(d.c_str(), 0755) != 0 && errno != EEXIST)
        throw CfError(CF_EINVAL, "cannot create directory " + d + ": " + strerror(errno));
}

std::string extent_str(int dim, const int n[3])
{
    char b[128];
    snprintf(b, sizeof b, "0 %d 0 %d 0 %d", n[0], n[1], dim == 3 ? n[2] : 0);
    return b;
}

std::string geo_attrs(int dim, const int n[3], const double origin[3], double dx)
{
    char b[512];
    snprintf(b, sizeof b, "WholeExtent=\"%s\" Origin=\"%.17g %.17g %.17g\" Spacing=\"%.17g %.17g %.17g\"",
             extent_str(dim, n).c_str(), origin[0], origin[1], dim == 3 ? origin[2] : 0.0, dx, dx,
             dx);
    return b;
}

} // namespace

// One ImageData piece from host arrays (q: n0 n1 n2 values; vel: dim arrays
// of as many, back to back; both x fastest).
void write_vti(const std::string& path, int dim, const int n[3], const double origin[3], double dx,
               int cycle, double time, double dt, const double* q, const double* vel)
{
    const size_t N = (size_t)n[0] * n[1] * (dim == 3 ? n[2] : 1);
    std::vector<double> v3(3 * N, 0.0);
    for (int d = 0; d < dim; ++d)
        for (size_t c = 0; c < N; ++c)
            v3[3 * c + d] = vel[d * N + c];
    const uint64_t qbytes = N * sizeof(double), vbytes = 3 * N * sizeof(double);
    FILE* f = fopen(path.c_str(), "wb");
    if (!f)
        throw CfError(CF_EINVAL, "cannot open " + path + ": " + strerror(errno));
    fprintf(f, "<?xml version=\"1.0\"?>\n"
               "<VTKFile type=\"ImageData\" version=\"1.0\" byte_order=\"LittleEndian\" "
               "header_type=\"UInt64\">\n");
    fprintf(f, "  <ImageData %s>\n", geo_attrs(dim, n, origin, dx).c_str());
    fprintf(f, "    <FieldData>\n"
               "      <DataArray type=\"Int32\" Name=\"CYCLE\" NumberOfTuples=\"1\" "
               "format=\"ascii\">%d</DataArray>\n"
               "      <DataArray type=\"Float64\" Name=\"TIME\" NumberOfTuples=\"1\" "
               "format=\"ascii\">%.17g</DataArray>\n"
               "      <DataArray type=\"Float64\" Name=\"DTIME\" NumberOfTuples=\"1\" "
               "format=\"ascii\">%.17g</DataArray>\n"
               "    </FieldData>\n",
            cycle, time, dt);
    fprintf(f, "    <Piece Extent=\"%s\">\n", extent_str(dim, n).c_str());
    fprintf(f, "      <CellData Scalars=\"quantity\" Vectors=\"velocity\">\n"
               "        <DataArray type=\"Float64\" Name=\"quantity\" NumberOfComponents=\"1\" "
               "format=\"appended\" offset=\"0\"/>\n"
               "        <DataArray type=\"Float64\" Name=\"velocity\" NumberOfComponents=\"3\" "
               "format=\"appended\" offset=\"%llu\"/>\n"
               "      </CellData>\n"
               "    </Piece>\n"
               "  </ImageData>\n"
               "  <AppendedData encoding=\"raw\">\n   _",
            (unsigned long long)(sizeof(uint64_t) + qbytes));
    bool ok = fwrite(&qbytes, sizeof qbytes, 1, f) == 1 && fwrite(q, 1, qbytes, f) == qbytes &&
              fwrite(&vbytes, sizeof vbytes, 1, f) == 1 &&
              fwrite(v3.data(), 1, vbytes, f) == vbytes;
    ok = fprintf(f, "\n  </AppendedData>\n</VTKFile>\n") > 0 && ok;
    ok = fclose(f) == 0 && ok;
    if (!ok)
        throw CfError(CF_EINVAL, "write to " + path + " failed");
}

// SiloWriter::siloWrite (SiloWriter.hpp:352-418) on one rank: the piece of
// group rank 0 and the master naming it.
void write_output(Solver& s, const std::string& dir, int step)
{
    const Geo& g = s.geo();
    const int n[3] = {g.n[0], g.n[1], g.dim == 3 ? g.n[2] : 1};
    const size_t N = (size_t)n[0] * n[1] * n[2];
    std::vector<double> q(N), vel((size_t)g.dim * N);
    s.cellOutput(q.data(), vel.data());
    double origin[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < g.dim; ++d)
        origin[d] = g.low[d] + (double)g.h * g.dx; // node 0 of the owned cells
    make_dir(dir);
    const std::string root = dir.empty() ? std::string(".") : dir;
    make_dir(root + "/raw");
    char piece[64], master[64];
    snprintf(piece, sizeof piece, "raw/CajitaFluidsOutput%05d%05d.vti", 0, step);
    snprintf(master, sizeof master, "CajitaFluids%05d.pvti", step);
    write_vti(root + "/" + piece, g.dim, n, origin, g.dx, step, s.time(), g.dt, q.data(),
              vel.data());
    const std::string mpath = root + "/" + master;
    FILE* f = fopen(mpath.c_str(), "w");
    if (!f)
        throw CfError(CF_EINVAL, "cannot open " + mpath + ": " + strerror(errno));
    fprintf(f, "<?xml version=\"1.0\"?>\n"
               "<VTKFile type=\"PImageData\" version=\"1.0\" byte_order=\"LittleEndian\" "
               "header_type=\"UInt64\">\n"
               "  <PImageData %s GhostLevel=\"0\">\n"
               "    <PCellData Scalars=\"quantity\" Vectors=\"velocity\">\n"
               "      <PDataArray type=\"Float64\" Name=\"quantity\" NumberOfComponents=\"1\"/>\n"
               "      <PDataArray type=\"Float64\" Name=\"velocity\" NumberOfComponents=\"3\"/>\n"
               "    </PCellData>\n"
               "    <Piece Extent=\"%s\" Source=\"%s\"/>\n"
               "  </PImageData>\n"
               "</VTKFile>\n",
            geo_attrs(g.dim, n, origin, g.dx).c_str(), extent_str(g.dim, n).c_str(), piece);
    if (fclose(f) != 0)
        throw CfError(CF_EINVAL, "write to " + mpath + " failed");
}

} // namespace cfmi
