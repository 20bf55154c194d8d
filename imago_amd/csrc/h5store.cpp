// h5store.cpp - the reference's SCF integral file, written and read with
// the HDF5 C library (host code; libimago_h5.so, separate from the GPU
// library so that it alone depends on HDF5).
//
// Layout (src/imago/hdf5SCF.F90:60-230, hdf5SCFIntg.F90:105-525): in the
// file, a group per k-point set; under it "atomIntgGroup" holding the
// groups atomOverlap, atomKEOverlap, atomMVOverlap, atomNPOverlap (one
// dataset per k-point, "0000001" ...), atomDMOverlap, atomMMOverlap,
// atomKOverlap, atomKOverlapPlusG (three subgroups "0000001" .. "0000003"
// with a dataset per k-point; the last two hold full n x n "real" and
// "imag" ones) and atomPotOverlap (a subgroup per potential term with a
// dataset per k-point).  The eigenvector, eigenvalue and potential groups
// the reference keeps in the same file are not written here.  Every
// dataset of a packed matrix has the Fortran shape
// (numComponents, valeDim (valeDim + 1) / 2) - in C (packed, components),
// the same bytes - chunked (the whole matrix up to 250e6 elements) and
// deflated at level 1; every group carries an integer "status" attribute,
// 0 when created and 1 once its matrices are complete, which a later run
// reads to skip what is done (integrals.F90:265-294).
#include <hdf5.h>

#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/imago_h5.h"

struct imago_h5 {
    hid_t file = -1, kgroup = -1, intg = -1;
    int nk = 0, n = 0, pot_dim = 0, ncomp = 2;
    hsize_t npacked = 0;
};

namespace {

thread_local std::string g_h5_err;

int h5fail(const char *what, const std::string &name)
{
    g_h5_err = std::string(what) + " " + name;
    return IMAGO_H5_E;
}

std::string seq_name(int i)   // Fortran "(i7.7)" of the 1-based index
{
    char b[16];
    std::snprintf(b, sizeof b, "%07d", i);
    return b;
}

hid_t open_or_create_group(hid_t parent, const std::string &name, bool create)
{
    if (create) return H5Gcreate2(parent, name.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    return H5Gopen2(parent, name.c_str(), H5P_DEFAULT);
}

int set_status_attr(hid_t group, int value, bool create)
{
    hid_t a;
    if (create) {
        const hsize_t one = 1;
        hid_t sp = H5Screate_simple(1, &one, nullptr);
        a = H5Acreate2(group, "status", H5T_NATIVE_INT, sp, H5P_DEFAULT, H5P_DEFAULT);
        H5Sclose(sp);
    } else {
        a = H5Aopen(group, "status", H5P_DEFAULT);
    }
    if (a < 0) return IMAGO_H5_E;
    const herr_t e = H5Awrite(a, H5T_NATIVE_INT, &value);
    H5Aclose(a);
    return e < 0 ? IMAGO_H5_E : IMAGO_H5_OK;
}

// dataset creation properties: chunks of at most 250e6 doubles, deflate
// level 1 (hdf5SCFIntg.F90:142-185, 322-343).  A packed matrix keeps whole
// component rows in a chunk; a full n x n one scales both sides alike.
hid_t chunked_plist(const hsize_t chunk[2])
{
    hid_t pl = H5Pcreate(H5P_DATASET_CREATE);
    H5Pset_layout(pl, H5D_CHUNKED);
    H5Pset_chunk(pl, 2, chunk);
    H5Pset_deflate(pl, 1);
    return pl;
}

hid_t packed_plist(hsize_t ncomp, hsize_t npacked)
{
    hsize_t chunk[2] = {npacked, ncomp};
    if (ncomp * npacked > 250000000ull) chunk[0] = 250000000ull / ncomp;
    return chunked_plist(chunk);
}

hid_t full_plist(hsize_t n)
{
    hsize_t chunk[2] = {n, n};
    const double nn = (double)n * (double)n;
    if (nn > 250000000.0) {
        const double q = std::sqrt(nn / 250000000.0);
        chunk[0] = chunk[1] = (hsize_t)((double)n / q);
    }
    return chunked_plist(chunk);
}

int create_datasets(hid_t group, const std::vector<std::string> &names, hid_t space, hid_t pl)
{
    for (const std::string &nm : names) {
        hid_t d = H5Dcreate2(group, nm.c_str(), H5T_NATIVE_DOUBLE, space, H5P_DEFAULT, pl,
                             H5P_DEFAULT);
        if (d < 0) return h5fail("cannot create dataset", nm);
        H5Dclose(d);
    }
    return IMAGO_H5_OK;
}

// the group of (kind, term) below atomIntgGroup
std::string group_path(int kind, int term)
{
    switch (kind) {
    case IMAGO_H5_OL: return "atomOverlap";
    case IMAGO_H5_KE: return "atomKEOverlap";
    case IMAGO_H5_NP: return "atomNPOverlap";
    case IMAGO_H5_MV: return "atomMVOverlap";
    case IMAGO_H5_EP: return "atomPotOverlap/" + seq_name(term + 1);
    default: return "";
    }
}

}   // namespace

extern "C" const char *imago_h5_last_error(void) { return g_h5_err.c_str(); }

extern "C" int imago_h5_open(const char *path, const char *kpoint_group, int num_kpoints,
                             int vale_dim, int pot_dim, int gamma, imago_h5 **out, int *created)
{
    if (!path || !kpoint_group || !out || num_kpoints < 1 || vale_dim < 1 || pot_dim < 0)
        return h5fail("bad arguments for", path ? path : "(null)");
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    imago_h5 *f = new imago_h5;
    f->nk = num_kpoints;
    f->n = vale_dim;
    f->pot_dim = pot_dim;
    f->ncomp = gamma ? 1 : 2;
    f->npacked = (hsize_t)vale_dim * (vale_dim + 1) / 2;
    // file access without a raw-data chunk cache (hdf5SCF.F90:108-115)
    hid_t fapl = H5Pcreate(H5P_FILE_ACCESS);
    int mdc = 0;
    size_t rdcc_n = 0, rdcc_b = 0;
    double w0 = 0.0;
    H5Pget_cache(fapl, &mdc, &rdcc_n, &rdcc_b, &w0);
    H5Pset_cache(fapl, mdc, 0, 0, w0);
    bool fresh_file = false;
    f->file = H5Fopen(path, H5F_ACC_RDWR, fapl);
    if (f->file < 0) {
        f->file = H5Fcreate(path, H5F_ACC_EXCL, H5P_DEFAULT, fapl);
        fresh_file = true;
    }
    H5Pclose(fapl);
    if (f->file < 0) {
        delete f;
        return h5fail("cannot open or create", path);
    }
    const bool fresh = fresh_file || H5Lexists(f->file, kpoint_group, H5P_DEFAULT) <= 0;
    f->kgroup = open_or_create_group(f->file, kpoint_group, fresh);
    if (f->kgroup >= 0) f->intg = open_or_create_group(f->kgroup, "atomIntgGroup", fresh);
    if (f->kgroup < 0 || f->intg < 0) {
        imago_h5_close(f);
        return h5fail("cannot open the integral groups of", kpoint_group);
    }
    if (fresh) {
        std::vector<std::string> kname;
        for (int k = 0; k < num_kpoints; ++k) kname.push_back(seq_name(k + 1));
        std::vector<std::string> riname;
        for (int k = 0; k < num_kpoints; ++k) {
            riname.push_back("real" + seq_name(k + 1));
            riname.push_back("imag" + seq_name(k + 1));
        }
        const hsize_t dims[2] = {f->npacked, (hsize_t)f->ncomp};
        hid_t space = H5Screate_simple(2, dims, nullptr);
        hid_t pl = packed_plist(f->ncomp, f->npacked);
        // the momentum-matrix K overlaps are full n x n real and imaginary
        // parts (fullVV, hdf5SCFIntg.F90:136-137, 314)
        const hsize_t fdims[2] = {(hsize_t)vale_dim, (hsize_t)vale_dim};
        hid_t fspace = H5Screate_simple(2, fdims, nullptr);
        hid_t fpl = full_plist((hsize_t)vale_dim);
        int rc = IMAGO_H5_OK;
        for (const char *g : {"atomOverlap", "atomKEOverlap", "atomMVOverlap", "atomNPOverlap"}) {
            hid_t gid = open_or_create_group(f->intg, g, true);
            if (gid < 0) rc = h5fail("cannot create group", g);
            if (rc == IMAGO_H5_OK) rc = create_datasets(gid, kname, space, pl);
            if (rc == IMAGO_H5_OK) rc = set_status_attr(gid, 0, true);
            if (gid >= 0) H5Gclose(gid);
        }
        for (const char *g : {"atomDMOverlap", "atomMMOverlap", "atomKOverlap", "atomKOverlapPlusG"}) {
            if (rc != IMAGO_H5_OK) break;
            hid_t gid = open_or_create_group(f->intg, g, true);
            if (gid < 0) {
                rc = h5fail("cannot create group", g);
                break;
            }
            const bool ri = std::string(g) == "atomKOverlap" || std::string(g) == "atomKOverlapPlusG";
            for (int x = 1; x <= 3 && rc == IMAGO_H5_OK; ++x) {
                hid_t sg = open_or_create_group(gid, seq_name(x), true);
                if (sg < 0) rc = h5fail("cannot create group", seq_name(x));
                if (rc == IMAGO_H5_OK)
                    rc = ri ? create_datasets(sg, riname, fspace, fpl)
                            : create_datasets(sg, kname, space, pl);
                if (sg >= 0) H5Gclose(sg);
            }
            if (rc == IMAGO_H5_OK) rc = set_status_attr(gid, 0, true);
            H5Gclose(gid);
        }
        hid_t pg = rc == IMAGO_H5_OK ? open_or_create_group(f->intg, "atomPotOverlap", true) : -1;
        if (rc == IMAGO_H5_OK && pg < 0) rc = h5fail("cannot create group", "atomPotOverlap");
        for (int j = 0; j < pot_dim && rc == IMAGO_H5_OK; ++j) {
            hid_t tg = open_or_create_group(pg, seq_name(j + 1), true);
            if (tg < 0) rc = h5fail("cannot create group", seq_name(j + 1));
            if (rc == IMAGO_H5_OK) rc = create_datasets(tg, kname, space, pl);
            if (rc == IMAGO_H5_OK) rc = set_status_attr(tg, 0, true);
            if (tg >= 0) H5Gclose(tg);
        }
        if (pg >= 0) H5Gclose(pg);
        H5Pclose(pl);
        H5Sclose(space);
        H5Pclose(fpl);
        H5Sclose(fspace);
        if (rc != IMAGO_H5_OK) {
            imago_h5_close(f);
            return rc;
        }
        H5Fflush(f->file, H5F_SCOPE_GLOBAL);
    }
    if (created) *created = fresh ? 1 : 0;
    *out = f;
    return IMAGO_H5_OK;
}

extern "C" int imago_h5_close(imago_h5 *f)
{
    if (!f) return IMAGO_H5_OK;
    if (f->intg >= 0) H5Gclose(f->intg);
    if (f->kgroup >= 0) H5Gclose(f->kgroup);
    if (f->file >= 0) {
        H5Fflush(f->file, H5F_SCOPE_GLOBAL);
        H5Fclose(f->file);
    }
    delete f;
    return IMAGO_H5_OK;
}

static int packed_io(imago_h5 *f, int kind, int term, int kpoint, double *buf, bool write)
{
    if (!f || !buf) return h5fail("null argument", "");
    const std::string gp = group_path(kind, term);
    if (gp.empty() || kpoint < 0 || kpoint >= f->nk ||
        (kind == IMAGO_H5_EP && (term < 0 || term >= f->pot_dim)))
        return h5fail("no such matrix in", gp.empty() ? "atomIntgGroup" : gp);
    const std::string path = gp + "/" + seq_name(kpoint + 1);
    hid_t d = H5Dopen2(f->intg, path.c_str(), H5P_DEFAULT);
    if (d < 0) return h5fail("cannot open dataset", path);
    hid_t sp = H5Dget_space(d);
    hsize_t dims[2] = {0, 0};
    const int rank = H5Sget_simple_extent_dims(sp, dims, nullptr);
    H5Sclose(sp);
    if (rank != 2 || dims[0] != f->npacked || dims[1] != (hsize_t)f->ncomp) {
        H5Dclose(d);
        return h5fail("unexpected shape of", path);
    }
    const herr_t e = write ? H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf)
                           : H5Dread(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf);
    H5Dclose(d);
    return e < 0 ? h5fail(write ? "cannot write" : "cannot read", path) : IMAGO_H5_OK;
}

extern "C" int imago_h5_write_packed(imago_h5 *f, int kind, int term, int kpoint,
                                     const double *packed)
{
    return packed_io(f, kind, term, kpoint, const_cast<double *>(packed), true);
}

extern "C" int imago_h5_read_packed(imago_h5 *f, int kind, int term, int kpoint, double *packed)
{
    return packed_io(f, kind, term, kpoint, packed, false);
}

static int status_io(imago_h5 *f, int kind, int term, int *value, bool write)
{
    if (!f || !value) return h5fail("null argument", "");
    const std::string gp = group_path(kind, term);
    if (gp.empty() || (kind == IMAGO_H5_EP && (term < 0 || term >= f->pot_dim)))
        return h5fail("no such group in", "atomIntgGroup");
    hid_t g = H5Gopen2(f->intg, gp.c_str(), H5P_DEFAULT);
    if (g < 0) return h5fail("cannot open group", gp);
    int rc = IMAGO_H5_OK;
    if (write) {
        rc = set_status_attr(g, *value, false);
        if (rc == IMAGO_H5_OK) H5Fflush(f->file, H5F_SCOPE_GLOBAL);
    } else {
        hid_t a = H5Aopen(g, "status", H5P_DEFAULT);
        if (a < 0 || H5Aread(a, H5T_NATIVE_INT, value) < 0) rc = h5fail("cannot read status of", gp);
        if (a >= 0) H5Aclose(a);
    }
    H5Gclose(g);
    return rc;
}

extern "C" int imago_h5_status(imago_h5 *f, int kind, int term, int *status)
{
    return status_io(f, kind, term, status, false);
}

extern "C" int imago_h5_set_status(imago_h5 *f, int kind, int term, int status)
{
    return status_io(f, kind, term, &status, true);
}
