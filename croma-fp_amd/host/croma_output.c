/*
 * croma_output.c -- spectra, emissivity and checkpoint files in the
 * reference's byte layouts (tracer_fp_output.c, tracer_fp_restart.c), written
 * from host arrays or straight from a batch's device-resident state.
 *
 * Device state reaches the file through two pinned staging buffers: the
 * copy of chunk k + 1 runs on the library stream while chunk k is written
 * with pwrite, so the PCIe copy and the file system overlap.
 */
#ifndef _GNU_SOURCE
#define _GNU_SOURCE
#endif
#include <errno.h>
#include <fcntl.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/types.h>
#include <unistd.h>

#include "croma_host.h"

#define OUT_PATH 4096
#define OUT_MSUN 1.99e+33            /* CONSTANTS.c M_sun */
#define OUT_STAGE_ROWS_BYTES (32u << 20) /* staging chunk of the device writer */

struct CromaHipOutput {
    int fd;
    int mode;
    size_t ntracer, nbin, nslot;
    char path[OUT_PATH];
};

static int io_fail(const char *what, const char *path)
{
    ch_set_error("%s '%s': %s", what, path ? path : "", strerror(errno));
    return CROMA_HIP_EIO;
}

static int pwrite_full(int fd, const void *buf, size_t nbyte, off_t off)
{
    const char *p = (const char *)buf;
    while (nbyte > 0) {
        const ssize_t w = pwrite(fd, p, nbyte, off);
        if (w < 0) {
            if (errno == EINTR) continue;
            return -1;
        }
        if (w == 0) {
            errno = EIO;
            return -1;
        }
        p += w;
        off += w;
        nbyte -= (size_t)w;
    }
    return 0;
}

static int ensure_dir(const char *path)
{
    struct stat st;
    if (!path || !*path) return -1;
    if (stat(path, &st) == 0) return S_ISDIR(st.st_mode) ? 0 : -1;
    if (mkdir(path, 0755) == 0) return 0;
    return (stat(path, &st) == 0 && S_ISDIR(st.st_mode)) ? 0 : -1;
}

CromaHipOutput *croma_hip_output_open(const char *dir, const char *stem, int rank, int mode, int ntracer,
                                      int nbin, int nslot, int resume_existing)
{
    CromaHipOutput *o;
    size_t nbyte;
    struct stat st;
    if (!dir || !*dir || !stem || !*stem || rank < 0 || ntracer <= 0 || nbin <= 0 || nslot <= 0 ||
        (mode != CROMA_HIP_OUT_MAP && mode != CROMA_HIP_OUT_TILE)) {
        ch_set_error("invalid output arguments");
        return 0;
    }
    if (!resume_existing && ensure_dir(dir) != 0) {
        io_fail("output directory", dir);
        return 0;
    }
    o = (CromaHipOutput *)calloc(1, sizeof(*o));
    if (!o) return 0;
    o->mode = mode;
    o->ntracer = (size_t)ntracer;
    o->nbin = (size_t)nbin;
    o->nslot = (size_t)nslot;
    snprintf(o->path, sizeof(o->path), mode == CROMA_HIP_OUT_TILE ? "%s/%s_tile_core%02d.bin" : "%s/%s_core%02d.bin",
             dir, stem, rank);
    nbyte = o->ntracer * o->nbin * o->nslot * sizeof(double);
    o->fd = resume_existing ? open(o->path, O_RDWR) : open(o->path, O_RDWR | O_CREAT | O_TRUNC, 0644);
    if (o->fd < 0) goto fail;
    if (resume_existing) {
        if (fstat(o->fd, &st) != 0) goto fail;
        if ((size_t)st.st_size != nbyte) {
            errno = EINVAL;
            goto fail;
        }
    } else {
        int e;
        if (ftruncate(o->fd, (off_t)nbyte) != 0) goto fail;
        e = posix_fallocate(o->fd, 0, (off_t)nbyte);
        if (e != 0 && e != EOPNOTSUPP && e != EINVAL) {
            errno = e;
            goto fail;
        }
    }
    return o;
fail:
    io_fail(resume_existing ? "existing output (missing or of another size)" : "output file", o->path);
    if (o->fd >= 0) close(o->fd);
    free(o);
    return 0;
}

/* per-cm^3 factor of tracer i (output_volume_factor, tracer_fp_output.c:175-188) */
static double volume_factor(const double *rho, const double *mass, size_t i)
{
    if (!rho || !mass) return 1.0;
    if (!(rho[i] > 0.0) || !(mass[i] > 0.0)) return 0.0;
    return rho[i] / (mass[i] * OUT_MSUN);
}

/* rows [first, first + count) of slot from buf (already scaled) */
static int write_rows(CromaHipOutput *o, size_t slot, size_t first, size_t count, const double *buf)
{
    const size_t rowb = o->nbin * sizeof(double);
    if (o->mode == CROMA_HIP_OUT_TILE) {
        const off_t off = (off_t)(((slot * o->ntracer) + first) * rowb);
        return pwrite_full(o->fd, buf, count * rowb, off);
    }
    /* map layout: tracer rows are nslot * nbin long; one pwrite per tracer */
    for (size_t i = 0; i < count; i++) {
        const off_t off = (off_t)(((first + i) * o->nslot + slot) * rowb);
        if (pwrite_full(o->fd, buf + i * o->nbin, rowb, off) != 0) return -1;
    }
    return 0;
}

static void scale_rows(double *dst, const double *src, size_t first, size_t count, size_t nbin, const double *rho,
                       const double *mass)
{
    for (size_t i = 0; i < count; i++) {
        const double f = volume_factor(rho, mass, first + i);
        for (size_t j = 0; j < nbin; j++) dst[i * nbin + j] = src[i * nbin + j] * f;
    }
}

int croma_hip_output_write(CromaHipOutput *o, int slot, const double *rows, const double *rho_gcc,
                           const double *tracer_mass)
{
    const int per_cc = rho_gcc && tracer_mass;
    if (!o || !rows || slot < 0 || (size_t)slot >= o->nslot) return CROMA_HIP_EINVAL;
    if (!per_cc) {
        if (write_rows(o, (size_t)slot, 0, o->ntracer, rows) != 0) return io_fail("output write", o->path);
        return 0;
    }
    {
        const size_t chunk = 4096;
        double *tmp = (double *)malloc(chunk * o->nbin * sizeof(double));
        size_t i;
        if (!tmp) return CROMA_HIP_ENOMEM;
        for (i = 0; i < o->ntracer; i += chunk) {
            const size_t cnt = (o->ntracer - i < chunk) ? o->ntracer - i : chunk;
            scale_rows(tmp, rows + i * o->nbin, i, cnt, o->nbin, rho_gcc, tracer_mass);
            if (write_rows(o, (size_t)slot, i, cnt, tmp) != 0) {
                free(tmp);
                return io_fail("output write", o->path);
            }
        }
        free(tmp);
    }
    return 0;
}

/* device rows d[0..ntr * n) into slot of o, double-buffered through pinned memory */
static int write_device_rows(CromaHipOutput *o, int slot, const double *d, void *stream, const double *rho,
                             const double *mass)
{
    size_t rows_per = OUT_STAGE_ROWS_BYTES / (o->nbin * sizeof(double));
    double *stage[2] = {0, 0}, *tmp = 0;
    void *ev[2] = {0, 0};
    size_t first, nchunk, k;
    int rc = 0;
    if (rows_per < 1) rows_per = 1;
    if (rows_per > o->ntracer) rows_per = o->ntracer;
    nchunk = (o->ntracer + rows_per - 1) / rows_per;
    for (k = 0; k < 2; k++) {
        stage[k] = (double *)ck_host_alloc(rows_per * o->nbin * sizeof(double));
        ev[k] = ck_event_create();
        if (!stage[k] || !ev[k]) {
            ch_set_error("output staging: %s", ck_last_error());
            rc = CROMA_HIP_ENOMEM;
            goto out;
        }
    }
    if (rho && mass && !(tmp = (double *)malloc(rows_per * o->nbin * sizeof(double)))) {
        rc = CROMA_HIP_ENOMEM;
        goto out;
    }
    /* copy of chunk k + 1 is queued before chunk k is written */
    for (k = 0; k <= nchunk && !rc; k++) {
        if (k < nchunk) {
            first = k * rows_per;
            const size_t cnt = (o->ntracer - first < rows_per) ? o->ntracer - first : rows_per;
            if (ck_d2h(stage[k & 1], d + first * o->nbin, cnt * o->nbin * sizeof(double), stream) ||
                ck_event_record(ev[k & 1], stream)) {
                ch_set_error("output copy: %s", ck_last_error());
                rc = CROMA_HIP_EDEVICE;
                break;
            }
        }
        if (k > 0) {
            const size_t j = k - 1, f = j * rows_per;
            const size_t cnt = (o->ntracer - f < rows_per) ? o->ntracer - f : rows_per;
            const double *src = stage[j & 1];
            if (ck_event_sync(ev[j & 1])) {
                ch_set_error("output copy: %s", ck_last_error());
                rc = CROMA_HIP_EDEVICE;
                break;
            }
            if (tmp) {
                scale_rows(tmp, src, f, cnt, o->nbin, rho, mass);
                src = tmp;
            }
            if (write_rows(o, (size_t)slot, f, cnt, src) != 0) rc = io_fail("output write", o->path);
        }
    }
out:
    ck_stream_sync(stream);
    for (k = 0; k < 2; k++) {
        ck_host_free(stage[k]);
        ck_event_destroy(ev[k]);
    }
    free(tmp);
    return rc;
}

int croma_hip_batch_write_spectra(CromaHipBatch *b, CromaHipOutput *cre_out, CromaHipOutput *crp_out, int slot,
                                  const double *rho_gcc, const double *tracer_mass)
{
    ChBatchState st;
    int rc;
    const double *rho = (rho_gcc && tracer_mass) ? rho_gcc : 0, *mass = rho ? tracer_mass : 0;
    if (!cre_out || ch_batch_state(b, &st) != 0) return CROMA_HIP_EINVAL;
    if (cre_out->ntracer != (size_t)st.ntr || cre_out->nbin != (size_t)st.n || slot < 0 ||
        (size_t)slot >= cre_out->nslot ||
        (crp_out && (crp_out->ntracer != (size_t)st.ntr || crp_out->nbin != (size_t)st.n ||
                     (size_t)slot >= crp_out->nslot))) {
        ch_set_error("output file shape does not match the batch");
        return CROMA_HIP_EINVAL;
    }
    rc = write_device_rows(cre_out, slot, st.d_cre, st.stream, rho, mass);
    if (rc || !crp_out) return rc;
    if (st.d_crp) return write_device_rows(crp_out, slot, st.d_crp, st.stream, rho, mass);
    {
        /* electron-only batch: the proton rows are zero */
        double *z = (double *)calloc(crp_out->ntracer * crp_out->nbin, sizeof(double));
        if (!z) return CROMA_HIP_ENOMEM;
        rc = croma_hip_output_write(crp_out, slot, z, 0, 0);
        free(z);
        return rc;
    }
}

int croma_hip_output_close(CromaHipOutput *o)
{
    int rc = 0;
    if (!o) return 0;
    if (o->fd >= 0) {
        if (fsync(o->fd) != 0) rc = io_fail("output sync", o->path);
        if (close(o->fd) != 0 && !rc) rc = io_fail("output close", o->path);
    }
    free(o);
    return rc;
}

int croma_hip_write_tracer_ids(const char *dir, int rank, int ntracer, const int *ids)
{
    char path[OUT_PATH];
    FILE *fp;
    int i;
    if (!dir || !*dir || rank < 0 || ntracer < 0 || (ntracer > 0 && !ids)) return CROMA_HIP_EINVAL;
    if (ensure_dir(dir) != 0) return io_fail("output directory", dir);
    snprintf(path, sizeof(path), "%s/tracerid_core%02d.txt", dir, rank);
    fp = fopen(path, "wb");
    if (!fp) return io_fail("tracer id file", path);
    for (i = 0; i < ntracer; i++)
        if (fprintf(fp, "%09d\n", ids[i]) < 0) {
            fclose(fp);
            return io_fail("tracer id file", path);
        }
    if (fclose(fp) != 0) return io_fail("tracer id file", path);
    return 0;
}

/* ------------------------------------------------------------------ */
/* TFPCHK1                                                             */
typedef struct {
    char magic[16];
    int version;
    int mpi_rank;
    int world_size;
    int ntracer;
    int ntracer_global;
    int nsnap;
    int nfreq;
    int np_bins;
    int npe_bins;
    int nsnp_i_value;
    int nsnp_f_value;
    int completed_snapshot;
    long tracer_start;
} CkptHeader;

_Static_assert(sizeof(CkptHeader) == 72, "checkpoint header is 72 bytes");

static void ckpt_path(char *path, size_t len, const char *dir, int rank)
{
    snprintf(path, len, "%s/checkpoint_rank%02d.bin", dir, rank);
}

static int ckpt_meta_ok(const CromaHipCheckpointMeta *m)
{
    return m && m->ntracer > 0 && m->np_bins > 0 && m->npe_bins > 0 && m->mpi_rank >= 0;
}

/* writes header + ids + the two state blocks produced by fill (host
 * arrays, or chunks copied from the device) to .tmp, syncs, renames */
typedef int (*ckpt_block_fn)(void *arg, int which, FILE *fp);

static int ckpt_write_common(const char *dir, const CromaHipCheckpointMeta *m, const int *ids, ckpt_block_fn fill,
                             void *arg)
{
    char path[OUT_PATH], tmp[OUT_PATH + 8];
    CkptHeader h;
    FILE *fp;
    if (!dir || !*dir || !ckpt_meta_ok(m) || !ids) return CROMA_HIP_EINVAL;
    if (ensure_dir(dir) != 0) return io_fail("checkpoint directory", dir);
    ckpt_path(path, sizeof(path), dir, m->mpi_rank);
    snprintf(tmp, sizeof(tmp), "%s.tmp", path);
    memset(&h, 0, sizeof(h));
    memcpy(h.magic, "TFPCHK1", 7);
    h.version = 1;
    h.mpi_rank = m->mpi_rank;
    h.world_size = m->world_size;
    h.ntracer = m->ntracer;
    h.ntracer_global = m->ntracer_global;
    h.nsnap = m->nsnap;
    h.nfreq = m->nfreq;
    h.np_bins = m->np_bins;
    h.npe_bins = m->npe_bins;
    h.nsnp_i_value = m->nsnp_i;
    h.nsnp_f_value = m->nsnp_f;
    h.completed_snapshot = m->completed_snapshot;
    h.tracer_start = m->tracer_start;
    fp = fopen(tmp, "wb");
    if (!fp) return io_fail("checkpoint", tmp);
    if (fwrite(&h, sizeof(h), 1, fp) != 1 || fwrite(ids, sizeof(int), (size_t)m->ntracer, fp) != (size_t)m->ntracer ||
        fill(arg, 0, fp) != 0 || fill(arg, 1, fp) != 0 || fflush(fp) != 0 || fsync(fileno(fp)) != 0) {
        const int e = errno;
        fclose(fp);
        unlink(tmp);
        errno = e;
        return io_fail("checkpoint write", tmp);
    }
    if (fclose(fp) != 0) return io_fail("checkpoint close", tmp);
    if (rename(tmp, path) != 0) return io_fail("checkpoint rename", path);
    return 0;
}

typedef struct {
    const double *crp, *cre;
    size_t ncrp, ncre;
} HostBlocks;

static int host_block(void *arg, int which, FILE *fp)
{
    const HostBlocks *hb = (const HostBlocks *)arg;
    const double *p = which == 0 ? hb->crp : hb->cre;
    const size_t cnt = which == 0 ? hb->ncrp : hb->ncre;
    return fwrite(p, sizeof(double), cnt, fp) == cnt ? 0 : -1;
}

int croma_hip_checkpoint_write(const char *dir, const CromaHipCheckpointMeta *meta, const int *ids,
                               const double *crp, const double *cre)
{
    HostBlocks hb;
    if (!ckpt_meta_ok(meta) || !crp || !cre) return CROMA_HIP_EINVAL;
    hb.crp = crp;
    hb.cre = cre;
    hb.ncrp = (size_t)meta->ntracer * meta->np_bins;
    hb.ncre = (size_t)meta->ntracer * meta->npe_bins;
    return ckpt_write_common(dir, meta, ids, host_block, &hb);
}

/* opens the checkpoint, checks header and tracer order; leaves fp at the crp block */
static int ckpt_open_checked(const char *dir, const CromaHipCheckpointMeta *m, const int *ids, FILE **fp_out,
                             int *completed)
{
    char path[OUT_PATH];
    CkptHeader h;
    FILE *fp;
    int *fids = 0, rc = CROMA_HIP_EIO;
    if (!dir || !*dir || !ckpt_meta_ok(m) || !ids || !completed) return CROMA_HIP_EINVAL;
    ckpt_path(path, sizeof(path), dir, m->mpi_rank);
    fp = fopen(path, "rb");
    if (!fp) return io_fail("checkpoint", path);
    if (fread(&h, sizeof(h), 1, fp) != 1 || strncmp(h.magic, "TFPCHK1", sizeof(h.magic)) != 0 || h.version != 1) {
        ch_set_error("checkpoint '%s': not a TFPCHK1 version 1 file", path);
        goto fail;
    }
    if (h.mpi_rank != m->mpi_rank || h.world_size != m->world_size || h.ntracer != m->ntracer ||
        h.ntracer_global != m->ntracer_global || h.nsnap != m->nsnap || h.nfreq != m->nfreq ||
        h.np_bins != m->np_bins || h.npe_bins != m->npe_bins || h.nsnp_i_value != m->nsnp_i ||
        h.nsnp_f_value != m->nsnp_f || h.tracer_start != m->tracer_start || h.completed_snapshot < 0 ||
        h.completed_snapshot >= m->nsnap) {
        ch_set_error("checkpoint metadata mismatch in '%s'", path);
        goto fail;
    }
    fids = (int *)malloc((size_t)m->ntracer * sizeof(int));
    if (!fids || fread(fids, sizeof(int), (size_t)m->ntracer, fp) != (size_t)m->ntracer) {
        ch_set_error("checkpoint '%s': truncated", path);
        goto fail;
    }
    if (memcmp(fids, ids, (size_t)m->ntracer * sizeof(int)) != 0) {
        ch_set_error("checkpoint tracer order mismatch in '%s'", path);
        goto fail;
    }
    free(fids);
    *completed = h.completed_snapshot;
    *fp_out = fp;
    return 0;
fail:
    free(fids);
    fclose(fp);
    return rc;
}

int croma_hip_checkpoint_read(const char *dir, const CromaHipCheckpointMeta *meta, const int *ids, double *crp,
                              double *cre, int *completed_snapshot)
{
    FILE *fp = 0;
    int done = 0, rc;
    size_t ncrp, ncre;
    if (!ckpt_meta_ok(meta) || !crp || !cre) return CROMA_HIP_EINVAL;
    rc = ckpt_open_checked(dir, meta, ids, &fp, &done);
    if (rc) return rc;
    ncrp = (size_t)meta->ntracer * meta->np_bins;
    ncre = (size_t)meta->ntracer * meta->npe_bins;
    if (fread(crp, sizeof(double), ncrp, fp) != ncrp || fread(cre, sizeof(double), ncre, fp) != ncre) {
        fclose(fp);
        ch_set_error("checkpoint truncated");
        return CROMA_HIP_EIO;
    }
    fclose(fp);
    *completed_snapshot = done;
    return 0;
}

/* device blocks: chunks of the state through one pinned buffer */
typedef struct {
    const double *d[2]; /* crp (NULL: zeros), cre */
    size_t count;       /* doubles per block */
    void *stream;
    double *stage;
    size_t stage_n;
} DevBlocks;

static int dev_block(void *arg, int which, FILE *fp)
{
    DevBlocks *db = (DevBlocks *)arg;
    size_t i;
    for (i = 0; i < db->count; i += db->stage_n) {
        const size_t cnt = (db->count - i < db->stage_n) ? db->count - i : db->stage_n;
        if (db->d[which]) {
            if (ck_d2h(db->stage, db->d[which] + i, cnt * sizeof(double), db->stream) || ck_stream_sync(db->stream))
                return -1;
        } else {
            memset(db->stage, 0, cnt * sizeof(double));
        }
        if (fwrite(db->stage, sizeof(double), cnt, fp) != cnt) return -1;
    }
    return 0;
}

int croma_hip_batch_checkpoint_write(CromaHipBatch *b, const char *dir, const CromaHipCheckpointMeta *meta,
                                     const int *ids)
{
    ChBatchState st;
    DevBlocks db;
    int rc;
    if (!ckpt_meta_ok(meta) || ch_batch_state(b, &st) != 0) return CROMA_HIP_EINVAL;
    if (meta->ntracer != st.ntr || meta->np_bins != st.n || meta->npe_bins != st.n) {
        ch_set_error("checkpoint shape does not match the batch");
        return CROMA_HIP_EINVAL;
    }
    if (ck_stream_sync(st.stream)) return CROMA_HIP_EDEVICE;
    db.d[0] = st.d_crp;
    db.d[1] = st.d_cre;
    db.count = (size_t)st.ntr * st.n;
    db.stream = st.stream;
    db.stage_n = OUT_STAGE_ROWS_BYTES / sizeof(double);
    if (db.stage_n > db.count) db.stage_n = db.count;
    db.stage = (double *)ck_host_alloc(db.stage_n * sizeof(double));
    if (!db.stage) return CROMA_HIP_ENOMEM;
    rc = ckpt_write_common(dir, meta, ids, dev_block, &db);
    ck_host_free(db.stage);
    return rc;
}

int croma_hip_batch_checkpoint_read(CromaHipBatch *b, const char *dir, const CromaHipCheckpointMeta *meta,
                                    const int *ids, int *completed_snapshot)
{
    ChBatchState st;
    FILE *fp = 0;
    double *stage;
    size_t count, stage_n, i;
    int which, rc, done = 0;
    if (!ckpt_meta_ok(meta) || ch_batch_state(b, &st) != 0 || !completed_snapshot) return CROMA_HIP_EINVAL;
    if (meta->ntracer != st.ntr || meta->np_bins != st.n || meta->npe_bins != st.n) {
        ch_set_error("checkpoint shape does not match the batch");
        return CROMA_HIP_EINVAL;
    }
    rc = ckpt_open_checked(dir, meta, ids, &fp, &done);
    if (rc) return rc;
    count = (size_t)st.ntr * st.n;
    stage_n = OUT_STAGE_ROWS_BYTES / sizeof(double);
    if (stage_n > count) stage_n = count;
    stage = (double *)ck_host_alloc(stage_n * sizeof(double));
    if (!stage) {
        fclose(fp);
        return CROMA_HIP_ENOMEM;
    }
    if (ck_stream_sync(st.stream)) rc = CROMA_HIP_EDEVICE;
    for (which = 0; which < 2 && !rc; which++) {
        double *d = which == 0 ? st.d_crp : st.d_cre;
        for (i = 0; i < count && !rc; i += stage_n) {
            const size_t cnt = (count - i < stage_n) ? count - i : stage_n;
            if (fread(stage, sizeof(double), cnt, fp) != cnt) {
                ch_set_error("checkpoint truncated");
                rc = CROMA_HIP_EIO;
            } else if (d && (ck_h2d(d + i, stage, cnt * sizeof(double), st.stream) || ck_stream_sync(st.stream))) {
                ch_set_error("checkpoint upload: %s", ck_last_error());
                rc = CROMA_HIP_EDEVICE;
            }
        }
    }
    fclose(fp);
    ck_host_free(stage);
    if (!rc) *completed_snapshot = done;
    return rc;
}
