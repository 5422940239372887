// lbm_output.hip -- the reference's macroscopic output file (OutputVTK,
// lbm.cpp:2290-2468) from the device state.
//
// k_vtk_pack writes every cell-data block of the file's appended section as
// Float32, in the file's cell order (i fastest, then j, then k = 1..Nz), into
// a device staging buffer; one copy on a second stream brings it to pinned
// host memory while the context's stream goes on stepping.  The host then
// writes the XML header with the reference's offsets, each block behind its
// UInt32 byte count, and the coordinate arrays.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "lbm_ctx.h"
#include "lbm_step.h"

using namespace lbm;

namespace {

constexpr int MAXB = 8;  // cell-data blocks

struct PackArgs {
  Geo g;
  int csize;             // 8: FP64 macro arrays, 4: FP32
  const void *rho, *ux, *uy, *uz;
  const void *T, *qx, *qy, *qz;
  const int16_t *type, *typeE;
  const uint8_t *mask;
  // force: mode 0 none, 1 per-cell arrays, 2 uniform; ibm: box override
  int force_mode;
  double Fu[3];
  const void *Fx, *Fy, *Fz;
  int ibm;
  BoxArgs box;
  double u_scale, T_a, T_b;
  int T_mode;
  int nblk;
  int kind[MAXB];        // 0 rho, 1 u, 2 F, 3 T, 4 q, 5 type, 6 typeE
  uint64_t off[MAXB];    // float offset of each block in the staging buffer
};

__device__ __forceinline__ double rd(const void *p, int csize, uint32_t i) {
  return csize == 8 ? ((const double *)p)[i] : (double)((const float *)p)[i];
}

// One thread per output cell c = ((k-1)*ny + j)*nx + i, k = 1..Nz.
__global__ __launch_bounds__(256) void k_vtk_pack(PackArgs a, float *__restrict__ out) {
  const uint32_t nx = a.g.nx, ny = a.g.ny, nzo = a.g.nz - 2;
  const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= (uint64_t)nx * ny * nzo) return;
  const uint32_t i = (uint32_t)(c % nx), j = (uint32_t)((c / nx) % ny), k = (uint32_t)(c / ((uint64_t)nx * ny)) + 1;
  const uint32_t d = (k * ny + j) * a.g.px + i;
  for (int b = 0; b < a.nblk; b++) {
    float *o = out + a.off[b];
    switch (a.kind[b]) {
      case 0:
        o[c] = (float)rd(a.rho, a.csize, d);
        break;
      case 1: {
        // val32 = (float)u / u_max: the float converts back to double for
        // the division (lbm.cpp:2451-2453)
        const double s = a.u_scale;
        o[3 * c + 0] = (float)((double)(float)rd(a.ux, a.csize, d) / s);
        o[3 * c + 1] = (float)((double)(float)rd(a.uy, a.csize, d) / s);
        o[3 * c + 2] = (float)((double)(float)rd(a.uz, a.csize, d) / s);
        break;
      }
      case 2: {
        double f[3] = {0.0, 0.0, 0.0};
        const int16_t t = a.type[d];
        const bool ibm_cell = t == 0 || t == 2 || t == 3;  // F, OUT, IN
        if (a.ibm && ibm_cell) {
          // IBLBM zeroes F of F/IN/OUT cells and spreads into the box
          // (lbm.cpp:1262-1285, 1347-1349)
          if (a.mask[d] & M_BOX) {
            const uint32_t bi = box_index(a.g, a.box, d);
            f[0] = a.box.Fx[bi];
            f[1] = a.box.Fy[bi];
            f[2] = a.box.Fz[bi];
          }
        } else if (a.force_mode == 1) {
          f[0] = rd(a.Fx, a.csize, d);
          f[1] = rd(a.Fy, a.csize, d);
          f[2] = rd(a.Fz, a.csize, d);
        } else if (a.force_mode == 2) {
          f[0] = a.Fu[0];
          f[1] = a.Fu[1];
          f[2] = a.Fu[2];
        }
        o[3 * c + 0] = (float)f[0];
        o[3 * c + 1] = (float)f[1];
        o[3 * c + 2] = (float)f[2];
        break;
      }
      case 3: {
        const double T = rd(a.T, a.csize, d);
        float v;
        if (a.T_mode == 1) v = (float)((double)(float)(T - a.T_a) / a.T_b);  // lbm.cpp:2586
        else if (a.T_mode == 2) v = (float)((double)(float)T / a.T_a - 1);   // lbm.cpp:2594
        else v = (float)T;
        o[c] = v;
        break;
      }
      case 4:
        o[3 * c + 0] = (float)rd(a.qx, a.csize, d);
        o[3 * c + 1] = (float)rd(a.qy, a.csize, d);
        o[3 * c + 2] = (float)rd(a.qz, a.csize, d);
        break;
      case 5:
        o[c] = (float)(int)a.type[d];
        break;
      case 6:
        o[c] = (float)(int)a.typeE[d];
        break;
    }
  }
}

}  // namespace

struct lbm_output {
  lbm_vtk_desc desc{};
  float *dev = nullptr;
  float *host = nullptr;  // pinned
  size_t cap = 0;         // floats
  hipStream_t copy = nullptr;
  hipEvent_t packed = nullptr, done = nullptr;
  bool pending = false;
  int nblk = 0;
  int kind[MAXB] = {};
  uint64_t off[MAXB] = {};
  int ncomp[MAXB] = {};
  size_t nfloat = 0;
};

void output_free(lbm_ctx *c) {
  lbm_output *o = c->out;
  if (!o) return;
  if (o->copy) (void)hipStreamSynchronize(o->copy);
  if (o->dev) (void)hipFree(o->dev);
  if (o->host) (void)hipHostFree(o->host);
  if (o->packed) (void)hipEventDestroy(o->packed);
  if (o->done) (void)hipEventDestroy(o->done);
  if (o->copy) (void)hipStreamDestroy(o->copy);
  delete o;
  c->out = nullptr;
}

namespace {

const char *kName[] = {"Density", "Velocity", "Force density", "Temperature", "Heat flux",
                       "CellType", "EnergyCellType"};

std::string block_name(const lbm_vtk_desc &d, int kind) {
  if (kind == 1) {
    if (d.velocity_name == 1) return "Velocity Ratio";
    if (d.velocity_name == 2) return "Velocity (Dimensional)";
  }
  if (kind == 3 && d.T_mode != 0) return "Temperature Ratio";
  return kName[kind];
}

int write_file(lbm_ctx *c, const char *path) {
  lbm_output *o = c->out;
  const lbm_vtk_desc &d = o->desc;
  const int Nx = (int)c->g.nx - 2, Ny = (int)c->g.ny - 2, zcomp = (int)c->g.nz - 2;
  FILE *fp = fopen(path, "wb");
  if (!fp) return LBM_ERR_ARG;
  const unsigned long ncell = (unsigned long)(Nx + 2) * (Ny + 2) * zcomp;
  unsigned long offset = 0;
  fprintf(fp, "<?xml version=\"1.0\"?>\n");
  fprintf(fp, "<VTKFile type=\"RectilinearGrid\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt32\">\n");
  fprintf(fp, "  <RectilinearGrid WholeExtent=\"0 %d 0 %d 0 %d\">\n", Nx + 2, Ny + 2, zcomp);
  fprintf(fp, "  <Piece Extent=\"0 %d 0 %d 0 %d\">\n", Nx + 2, Ny + 2, zcomp);
  fprintf(fp, "    <PointData>\n");
  fprintf(fp, "    </PointData>\n");
  fprintf(fp, "    <CellData>\n");
  for (int b = 0; b < o->nblk; b++) {
    const std::string nm = block_name(d, o->kind[b]);
    if (o->ncomp[b] == 3)
      fprintf(fp, "      <DataArray type=\"Float32\" Name=\"%s\" NumberOfComponents=\"3\" format=\"appended\" offset=\"%ld\" />\n",
              nm.c_str(), (long)offset);
    else
      fprintf(fp, "      <DataArray type=\"Float32\" Name=\"%s\" format=\"appended\" offset=\"%ld\" />\n", nm.c_str(),
              (long)offset);
    offset += 4 + (unsigned long)o->ncomp[b] * 4 * ncell;
  }
  fprintf(fp, "    </CellData>\n");
  fprintf(fp, "    <Coordinates>\n");
  fprintf(fp, "      <DataArray type=\"Float32\" Name=\"CoordinateX\" format=\"appended\" offset=\"%ld\" />\n", (long)offset);
  offset += 4 + 1 * 4 * (Nx + 3);
  fprintf(fp, "      <DataArray type=\"Float32\" Name=\"CoordinateY\" format=\"appended\" offset=\"%ld\" />\n", (long)offset);
  offset += 4 + 1 * 4 * (Ny + 3);
  fprintf(fp, "      <DataArray type=\"Float32\" Name=\"CoordinateZ\" format=\"appended\" offset=\"%ld\" />\n", (long)offset);
  offset += 4 + 1 * 4 * (zcomp + 1);
  fprintf(fp, "    </Coordinates>\n");
  fprintf(fp, "  </Piece>\n");
  fprintf(fp, "  </RectilinearGrid>\n");
  fprintf(fp, "  <AppendedData encoding=\"raw\">");
  fprintf(fp, "_");
  for (int b = 0; b < o->nblk; b++) {
    const unsigned int sz = (unsigned int)(o->ncomp[b] * 4 * ncell);
    fwrite(&sz, sizeof(unsigned int), 1, fp);
    fwrite(o->host + o->off[b], sizeof(float), (size_t)o->ncomp[b] * ncell, fp);
  }
  // coordinates (vertices), lbm.cpp:2617-2642
  const double dx = 1.0, dy = 1.0, dz = 1.0;
  unsigned int sz = 4 * (unsigned int)(Nx + 3);
  fwrite(&sz, sizeof(unsigned int), 1, fp);
  for (int i = 0; i < Nx + 3; i++) {
    const float v = (float)(i * dx);
    fwrite(&v, sizeof(float), 1, fp);
  }
  const double H = Ny + 2;
  sz = 4 * (unsigned int)(Ny + 3);
  fwrite(&sz, sizeof(unsigned int), 1, fp);
  for (int j = 0; j < Ny + 3; j++) {
    const float v = d.poiseuille_coords ? (float)((double)(float)(j * dy) / H * 2 - 1) : (float)(j * dy);
    fwrite(&v, sizeof(float), 1, fp);
  }
  sz = 4 * (unsigned int)(zcomp + 1);
  fwrite(&sz, sizeof(unsigned int), 1, fp);
  for (int k = 0; k < zcomp + 1; k++) {
    const float v = d.poiseuille_coords ? (float)((double)(float)(k * dz) / H * 2 - 1) : (float)(k * dz);
    fwrite(&v, sizeof(float), 1, fp);
  }
  fprintf(fp, "  </AppendedData>\n");
  fprintf(fp, "</VTKFile>\n");
  const bool ok = ferror(fp) == 0;
  fclose(fp);
  return ok ? LBM_OK : LBM_ERR_ARG;
}

}  // namespace

extern "C" {

int lbm_output_vtk_begin(lbm_ctx *c, const lbm_vtk_desc *d) {
  if (!c || !d) return LBM_ERR_ARG;
  if (c->state == ST_NONE) return LBM_ERR_STATE;
  if (d->thermal && !c->thermal) return LBM_ERR_STATE;
  if (!(d->u_scale != 0) || (d->T_mode == 1 && !(d->T_b != 0)) || (d->T_mode == 2 && !(d->T_a != 0)))
    return LBM_ERR_ARG;
  HIPCHK(hipSetDevice(c->dev));
  if (!c->out) c->out = new lbm_output();
  lbm_output *o = c->out;
  if (o->pending) {  // the previous output was never written: drop it
    HIPCHK(hipEventSynchronize(o->done));
    o->pending = false;
  }
  if (!o->copy) {
    HIPCHK(hipStreamCreateWithFlags(&o->copy, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&o->packed, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&o->done, hipEventDisableTiming));
  }
  o->desc = *d;
  // blocks in the file's order (lbm.cpp:2321-2375)
  int kinds[MAXB], ncomp[MAXB], nb = 0;
  kinds[nb] = 0; ncomp[nb++] = 1;
  kinds[nb] = 1; ncomp[nb++] = 3;
  if (d->force) { kinds[nb] = 2; ncomp[nb++] = 3; }
  if (d->thermal) {
    kinds[nb] = 3; ncomp[nb++] = 1;
    kinds[nb] = 4; ncomp[nb++] = 3;
  }
  kinds[nb] = 5; ncomp[nb++] = 1;
  if (d->thermal) { kinds[nb] = 6; ncomp[nb++] = 1; }
  const size_t ncell = (size_t)c->g.nx * c->g.ny * (c->g.nz - 2);
  size_t total = 0;
  PackArgs a{};
  for (int b = 0; b < nb; b++) {
    o->kind[b] = a.kind[b] = kinds[b];
    o->ncomp[b] = ncomp[b];
    o->off[b] = a.off[b] = total;
    total += (size_t)ncomp[b] * ncell;
  }
  o->nblk = a.nblk = nb;
  o->nfloat = total;
  if (total > o->cap) {
    if (o->dev) (void)hipFree(o->dev);
    if (o->host) (void)hipHostFree(o->host);
    o->dev = nullptr;
    o->host = nullptr;
    o->cap = 0;
    if (hipMalloc((void **)&o->dev, total * 4) != hipSuccess) return LBM_ERR_NOMEM;
    if (hipHostMalloc((void **)&o->host, total * 4, hipHostMallocDefault) != hipSuccess) return LBM_ERR_NOMEM;
    o->cap = total;
  }
  a.g = c->g;
  a.csize = (int)c->csize;
  a.rho = c->mac[0];
  a.ux = c->mac[1];
  a.uy = c->mac[2];
  a.uz = c->mac[3];
  a.T = c->T;
  a.qx = c->qv[0];
  a.qy = c->qv[1];
  a.qz = c->qv[2];
  a.type = c->type;
  a.typeE = c->typeE;
  a.mask = c->mask;
  a.force_mode = c->force_mode;
  for (int q = 0; q < 3; q++) a.Fu[q] = c->Fu[q];
  a.Fx = c->F[0];
  a.Fy = c->F[1];
  a.Fz = c->F[2];
  a.ibm = ibm_active(c) ? 1 : 0;
  a.box = c->box;
  a.u_scale = d->u_scale;
  a.T_a = d->T_a;
  a.T_b = d->T_b;
  a.T_mode = d->T_mode;
  hipLaunchKernelGGL(k_vtk_pack, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, c->stream, a, o->dev);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(o->packed, c->stream));
  HIPCHK(hipStreamWaitEvent(o->copy, o->packed, 0));
  HIPCHK(hipMemcpyAsync(o->host, o->dev, total * 4, hipMemcpyDeviceToHost, o->copy));
  HIPCHK(hipEventRecord(o->done, o->copy));
  o->pending = true;
  return LBM_OK;
}

int lbm_output_vtk_end(lbm_ctx *c, const char *path) {
  if (!c || !path) return LBM_ERR_ARG;
  lbm_output *o = c->out;
  if (!o || !o->pending) return LBM_ERR_STATE;
  HIPCHK(hipSetDevice(c->dev));
  HIPCHK(hipEventSynchronize(o->done));
  o->pending = false;
  return write_file(c, path);
}

int lbm_output_vtk(lbm_ctx *c, const lbm_vtk_desc *d, const char *path) {
  int e = lbm_output_vtk_begin(c, d);
  if (e != LBM_OK) return e;
  return lbm_output_vtk_end(c, path);
}

}  // extern "C"
