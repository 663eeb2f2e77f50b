/*
 * io.c -- on-disk formats on either side of the path (host side, C).
 *
 * Restart files: the reference writes the conserved state with
 * VecView(U, PetscViewerBinary) to restart1.bin / restart2.bin and reads it
 * back with VecLoad (post_process.c:47-63, hype.c:316-330).  PETSc's binary
 * Vec layout (a PETSc build with 32-bit indices): big-endian int32
 * VEC_FILE_CLASSID = 1211214, big-endian int32 length, then the values as
 * big-endian IEEE doubles.  A DMDA vector is written in natural ordering,
 * which on one rank is the [j][i][c] layout of the state.  The viewer also
 * writes a text file <name>.info holding "-vecload_block_size <dof>".
 *
 * Solution snapshots: MonitorFunction writes the primitive variables as a
 * VTK XML structured grid (DMView + VecView on a PetscViewerVTK,
 * post_process.c:29-41) that plot.py reads with vtkXMLStructuredGridReader
 * (plot.py:182-210): point coordinates from the DMDA coordinates and one
 * scalar point-data array per component.  The file written here follows the
 * VTK XML format (little-endian, UInt64 block headers, raw appended data).
 */
#define _GNU_SOURCE
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "common.h"

#define VEC_FILE_CLASSID 1211214

static void put_be32(unsigned char *b, unsigned v)
{
    b[0] = (unsigned char)(v >> 24);
    b[1] = (unsigned char)(v >> 16);
    b[2] = (unsigned char)(v >> 8);
    b[3] = (unsigned char)v;
}

static unsigned get_be32(const unsigned char *b)
{
    return ((unsigned)b[0] << 24) | ((unsigned)b[1] << 16) | ((unsigned)b[2] << 8) | (unsigned)b[3];
}

static void put_be64d(unsigned char *b, double d)
{
    unsigned long long u;
    int k;
    memcpy(&u, &d, sizeof(u));
    for (k = 0; k < 8; ++k) b[k] = (unsigned char)(u >> (56 - 8 * k));
}

static double get_be64d(const unsigned char *b)
{
    unsigned long long u = 0;
    double d;
    int k;
    for (k = 0; k < 8; ++k) u = (u << 8) | b[k];
    memcpy(&d, &u, sizeof(d));
    return d;
}

int hype_write_vec_binary(const char *path, const double *U, long n, int block_size)
{
    FILE *f;
    unsigned char hdr[8], buf[8 * 512];
    long i = 0;
    char info[4096];
    if (!path || !U || n < 0 || n > 0x7fffffffL) return HYPE_E_INVALID;
    f = fopen(path, "wb");
    if (!f) return HYPE_E_INVALID;
    put_be32(hdr, VEC_FILE_CLASSID);
    put_be32(hdr + 4, (unsigned)n);
    if (fwrite(hdr, 1, 8, f) != 8) goto fail;
    while (i < n) {
        long m = n - i < 512 ? n - i : 512, k;
        for (k = 0; k < m; ++k) put_be64d(buf + 8 * k, U[i + k]);
        if (fwrite(buf, 8, (size_t)m, f) != (size_t)m) goto fail;
        i += m;
    }
    if (fclose(f) != 0) return HYPE_E_INVALID;
    if (block_size > 1) {
        if (snprintf(info, sizeof(info), "%s.info", path) >= (int)sizeof(info)) return HYPE_E_INVALID;
        f = fopen(info, "w");
        if (!f) return HYPE_E_INVALID;
        fprintf(f, "-vecload_block_size %d\n", block_size);
        if (fclose(f) != 0) return HYPE_E_INVALID;
    }
    return HYPE_OK;
fail:
    fclose(f);
    return HYPE_E_INVALID;
}

int hype_read_vec_binary(const char *path, double *U, long n)
{
    FILE *f;
    unsigned char hdr[8], buf[8 * 512];
    long i = 0;
    if (!path || !U || n < 0) return HYPE_E_INVALID;
    f = fopen(path, "rb");
    if (!f) return HYPE_E_NOTFOUND;
    if (fread(hdr, 1, 8, f) != 8 || get_be32(hdr) != VEC_FILE_CLASSID || (long)get_be32(hdr + 4) != n) {
        fclose(f);
        return HYPE_E_INVALID;
    }
    while (i < n) {
        long m = n - i < 512 ? n - i : 512, k;
        if (fread(buf, 8, (size_t)m, f) != (size_t)m) {
            fclose(f);
            return HYPE_E_INVALID;
        }
        for (k = 0; k < m; ++k) U[i + k] = get_be64d(buf + 8 * k);
        i += m;
    }
    fclose(f);
    return HYPE_OK;
}

/* component names of the primitive state written by the reference's
 * ComputePrimitiveVariables (kapila_5eqn.c:14-35, euler.c:14-23,
 * baer_nunziato.c:14-41) */
static const char *prim_name(int system, int c)
{
    static const char *eu[] = {"rho", "u", "v", "p"};
    static const char *ka[] = {"rho_1", "rho_2", "u", "v", "p", "alpha_1"};
    static const char *bn[] = {"rho_1", "u_1", "v_1", "p_1", "rho_2", "u_2", "v_2", "p_2", "phi_1"};
    if (system == HYPE_EULER) return eu[c];
    if (system == HYPE_KAPILA) return ka[c];
    return bn[c];
}

int hype_write_vts(const char *path, const hype_problem *p, const double *W, int coord_mode)
{
    FILE *f;
    int nv, c, i, j;
    const long npts = p ? (long)p->nx * p->ny : 0;
    unsigned long long off = 0, nb;
    double *col;
    if (!path || !p || !W) return HYPE_E_INVALID;
    HYPE_TRY(hype_check_problem(p, &nv));
    f = fopen(path, "wb");
    if (!f) return HYPE_E_INVALID;
    fprintf(f, "<?xml version=\"1.0\"?>\n");
    fprintf(f, "<VTKFile type=\"StructuredGrid\" version=\"1.0\" byte_order=\"LittleEndian\" "
               "header_type=\"UInt64\">\n");
    fprintf(f, "  <StructuredGrid WholeExtent=\"0 %d 0 %d 0 0\">\n", p->nx - 1, p->ny - 1);
    fprintf(f, "    <Piece Extent=\"0 %d 0 %d 0 0\">\n", p->nx - 1, p->ny - 1);
    fprintf(f, "      <Points>\n        <DataArray type=\"Float64\" Name=\"Position\" "
               "NumberOfComponents=\"3\" format=\"appended\" offset=\"%llu\"/>\n      </Points>\n",
            off);
    off += 8 + (unsigned long long)npts * 3 * 8;
    fprintf(f, "      <PointData>\n");
    for (c = 0; c < nv; ++c) {
        fprintf(f, "        <DataArray type=\"Float64\" Name=\"%s\" format=\"appended\" offset=\"%llu\"/>\n",
                prim_name(p->system, c), off);
        off += 8 + (unsigned long long)npts * 8;
    }
    fprintf(f, "      </PointData>\n    </Piece>\n  </StructuredGrid>\n");
    fprintf(f, "  <AppendedData encoding=\"raw\">\n_");
    /* points (little-endian doubles: the host's byte order on x86-64) */
    nb = (unsigned long long)npts * 3 * 8;
    fwrite(&nb, 8, 1, f);
    for (j = 0; j < p->ny; ++j)
        for (i = 0; i < p->nx; ++i) {
            double xyz[3];
            xyz[0] = hype_cell_coord(p, 0, i, coord_mode);
            xyz[1] = hype_cell_coord(p, 1, j, coord_mode);
            xyz[2] = 0.0;
            fwrite(xyz, 8, 3, f);
        }
    col = (double *)malloc((size_t)npts * sizeof(double));
    if (!col) {
        fclose(f);
        return HYPE_E_NOMEM;
    }
    for (c = 0; c < nv; ++c) {
        long k;
        for (k = 0; k < npts; ++k) col[k] = W[k * nv + c];
        nb = (unsigned long long)npts * 8;
        fwrite(&nb, 8, 1, f);
        fwrite(col, 8, (size_t)npts, f);
    }
    free(col);
    fprintf(f, "\n  </AppendedData>\n</VTKFile>\n");
    if (ferror(f)) {
        fclose(f);
        return HYPE_E_INVALID;
    }
    return fclose(f) == 0 ? HYPE_OK : HYPE_E_INVALID;
}
