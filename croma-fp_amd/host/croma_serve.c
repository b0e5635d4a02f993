/*
 * croma_serve.c -- request loop of the tracer update (continuous batching).
 *
 * Chunks of tracers (any size, grid n in {64,128,256,512}, electron_only or
 * proton-seeded / two-species, any D_pp model) are submitted at any time.  A
 * scheduler thread admits them against a fixed HBM pool (first fit, in
 * arrival order), uploads their spectra, backgrounds of every snapshot and
 * sources once (they stay resident until the chunk finishes), builds their
 * IC cooling rows and their nsub schedule on the device, and
 * then advances every admitted chunk snapshot by snapshot.  Chunks of one
 * launch class (grid size x species set) share one persistent fused launch
 * per step: the task list of a launch is the cost-ordered union of the
 * tracers of every chunk of the class, each task advancing its tracer over
 * that chunk's next snapshot window.  Classes run on their own streams and
 * progress independently; a chunk that arrives while others are on the GPU
 * joins the next launch of its class.  Every tracer's arithmetic is the same
 * whether its chunk runs alone or in a mix (tracers never interact), so a
 * result does not depend on what else is being served.
 *
 * Latency of a chunk = submit -> final spectra copied to host memory.
 */
#include "croma_host.h"

#include <errno.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#define SV_NCLASS 8 /* 4 grid sizes x {electron_only, two-species} */
#define SV_ALIGN 256

enum { SV_QUEUED = 0, SV_ACTIVE, SV_RETIRING, SV_DONE, SV_FAILED, SV_CANCELLED };

typedef struct {
    long long id;
    int state, status, cancel;
    int ntr, n, nidx, cls, protons, nsnap;
    CkParams prm;
    double *z, *dt;    /* host [nsnap + 1], [nsnap] */
    double *bg;        /* host [(nsnap + 1)][8][ntr] */
    double *h_in;      /* pinned block holding the present inputs below */
    double *h_cre, *h_crp, *h_qpi, *h_qepri, *h_mass;
    double *h_out;     /* pinned: cre | crp (two-species) */
    int has_cre, has_crp, has_qpi, has_qepri, has_mass;
    int *h_sched;      /* pinned: nsub [nsnap][ntr] | n_on [nsnap][ntr] */
    double *h_zb;      /* pinned [(nsnap + 1)][3] */
    size_t bytes, off;
    double *d_crp, *d_cre, *d_qpi, *d_qepri, *d_mass, *d_bg, *d_ric, *d_zb;
    int *d_nsub, *d_non;
    void *ev_ready, *ev_done;
    int snap_next, inflight, ready;
    double elapsed;
    double t_submit, t_admit, t_done;
    long long cell_steps;
} SvReq;

typedef struct {
    size_t off, size;
} SvBlk;

typedef struct {
    void *stream, *ev;
    int busy;
    int *d_ctrl;         /* counter, err */
    int *h_err;          /* pinned */
    CkSlot *h_slots, *d_slots;
    CkTask *h_tasks, *d_tasks;
    int slot_cap, task_cap;
    SvReq **launched;
    int nlaunched, launched_cap;
    /* split launches (two species, no eta cap): segment offsets per launched
     * (chunk, snapshot) and the class's segment-row buffers */
    int *h_segoff, *d_segoff;
    long segoff_cap;
    double *d_segrec, *d_seginje;
    long segrows_cap;
} SvClass;

struct CromaHipServer {
    CromaHipParams uprm;
    double pmin, pmax, pemin, pemax;
    CromaHipContext *ctx[4];
    SvClass cls[SV_NCLASS];
    void *copy_stream;
    int *d_admit_err;
    char *d_pool;
    size_t pool_size, in_use, peak;
    SvBlk *fl;
    int nfl, capfl;
    /* scratch for task ordering */
    CkTask *tmp_tasks;
    int *keys;
    int tmp_cap;
    int hist[CH_HIST];
    /* requests by id (id = index + 1) */
    SvReq **reqs;
    long long nreq, capreq;
    long long *queue; /* pending ids, FIFO */
    long long qhead, qtail, qcap;
    SvReq **active;
    int nactive, capactive;
    pthread_t thr;
    int thr_started;
    pthread_mutex_t mu;
    pthread_cond_t cv_work, cv_done;
    int stop;
    CromaHipServeStats st;
    int head_age;
};

static const int SV_NS[4] = {64, 128, 256, 512};

static int sv_nidx(int n)
{
    int k;
    for (k = 0; k < 4; k++)
        if (SV_NS[k] == n) return k;
    return -1;
}

static size_t sv_round(size_t b) { return (b + SV_ALIGN - 1) & ~(size_t)(SV_ALIGN - 1); }

size_t croma_hip_chunk_bytes(int ntracer, int n, int species, int nsnap)
{
    const size_t st = (size_t)ntracer * n * sizeof(double);
    const int sp = (species == CROMA_HIP_SPECIES_ELECTRON) ? 1 : 2;
    size_t b = 0;
    if (ntracer <= 0 || n <= 0 || nsnap <= 0) return 0;
    b += sv_round(st) * sp;                                             /* state */
    b += sv_round(st) * sp;                                             /* sources */
    b += sv_round((size_t)ntracer * sizeof(double));                    /* mass */
    b += sv_round((size_t)(nsnap + 1) * CK_NBG * ntracer * sizeof(double)); /* backgrounds */
    b += 2 * sv_round((size_t)nsnap * ntracer * sizeof(int));           /* nsub, n_on */
    b += sv_round((size_t)(nsnap + 1) * n * sizeof(double));            /* IC rows */
    b += sv_round((size_t)(nsnap + 1) * 3 * sizeof(double));            /* CMB bounds */
    return b;
}

/* ---- first-fit pool allocator ---------------------------------------- */
static int sv_alloc(CromaHipServer *s, size_t size, size_t *off)
{
    int i;
    for (i = 0; i < s->nfl; i++)
        if (s->fl[i].size >= size) {
            *off = s->fl[i].off;
            s->fl[i].off += size;
            s->fl[i].size -= size;
            if (s->fl[i].size == 0) {
                memmove(&s->fl[i], &s->fl[i + 1], (size_t)(s->nfl - i - 1) * sizeof(SvBlk));
                s->nfl--;
            }
            s->in_use += size;
            if (s->in_use > s->peak) s->peak = s->in_use;
            return 0;
        }
    return -1;
}

static int sv_release(CromaHipServer *s, size_t off, size_t size)
{
    int i = 0;
    while (i < s->nfl && s->fl[i].off < off) i++;
    if (s->nfl == s->capfl) {
        SvBlk *nf = (SvBlk *)realloc(s->fl, (size_t)(s->capfl * 2 + 8) * sizeof(SvBlk));
        if (!nf) return -1;
        s->fl = nf;
        s->capfl = s->capfl * 2 + 8;
    }
    memmove(&s->fl[i + 1], &s->fl[i], (size_t)(s->nfl - i) * sizeof(SvBlk));
    s->fl[i].off = off;
    s->fl[i].size = size;
    s->nfl++;
    /* coalesce with the right, then the left neighbour */
    if (i + 1 < s->nfl && s->fl[i].off + s->fl[i].size == s->fl[i + 1].off) {
        s->fl[i].size += s->fl[i + 1].size;
        memmove(&s->fl[i + 1], &s->fl[i + 2], (size_t)(s->nfl - i - 2) * sizeof(SvBlk));
        s->nfl--;
    }
    if (i > 0 && s->fl[i - 1].off + s->fl[i - 1].size == s->fl[i].off) {
        s->fl[i - 1].size += s->fl[i].size;
        memmove(&s->fl[i], &s->fl[i + 1], (size_t)(s->nfl - i - 1) * sizeof(SvBlk));
        s->nfl--;
    }
    s->in_use -= size;
    return 0;
}

/* ---- request bookkeeping --------------------------------------------- */
/* State changes happen on the scheduler thread and are read by waiters
 * under the lock; the cancel flag is set by any thread. */
static void sv_set_state(CromaHipServer *s, SvReq *r, int state)
{
    pthread_mutex_lock(&s->mu);
    r->state = state;
    pthread_mutex_unlock(&s->mu);
}

static int sv_cancelled(const SvReq *r) { return __atomic_load_n(&r->cancel, __ATOMIC_ACQUIRE); }

static void sv_req_free_host(SvReq *r)
{
    free(r->z);
    free(r->dt);
    free(r->bg);
    ck_host_free(r->h_in);
    ck_host_free(r->h_out);
    ck_host_free(r->h_sched);
    ck_host_free(r->h_zb);
    ck_event_destroy(r->ev_ready);
    ck_event_destroy(r->ev_done);
    r->z = r->dt = r->bg = r->h_in = r->h_out = r->h_zb = 0;
    r->h_cre = r->h_crp = r->h_qpi = r->h_qepri = r->h_mass = 0;
    r->h_sched = 0;
    r->ev_ready = r->ev_done = 0;
}

static CromaHipContext *sv_ctx(CromaHipServer *s, int nidx)
{
    if (!s->ctx[nidx]) {
        CromaHipGridSpec gs;
        gs.n = SV_NS[nidx];
        gs.pmin = s->pmin;
        gs.pmax = s->pmax;
        gs.pemin = s->pemin;
        gs.pemax = s->pemax;
        s->ctx[nidx] = croma_hip_context_create(&gs, &s->uprm);
    }
    return s->ctx[nidx];
}

/* Tracer-substeps of a request, from its nsub schedule once that is on the host. */
static void sv_count_steps(SvReq *r)
{
    const size_t cnt = (size_t)r->nsnap * r->ntr;
    long long t = 0;
    size_t i;
    for (i = 0; i < cnt; i++) t += r->h_sched[i];
    r->cell_steps = t;
}

/* Host preparation and uploads of an admitted request (scheduler thread). */
static int sv_prepare(CromaHipServer *s, SvReq *r)
{
    CromaHipContext *c = sv_ctx(s, r->nidx);
    const ChGrid *g;
    const size_t m = (size_t)r->ntr * r->n;
    char *p;
    int q;
    if (!c) return CROMA_HIP_EDEVICE;
    g = ch_ctx_grid(c);
    if (r->protons && ch_ctx_fqe(c)) return CROMA_HIP_EDEVICE;
    /* device layout inside the request's pool block */
    p = s->d_pool + r->off;
#define SV_TAKE(ptr, type, bytes)                                                                                      \
    do {                                                                                                               \
        ptr = (type *)p;                                                                                               \
        p += sv_round(bytes);                                                                                          \
    } while (0)
    SV_TAKE(r->d_cre, double, m * sizeof(double));
    if (r->protons) SV_TAKE(r->d_crp, double, m * sizeof(double));
    if (r->has_qepri) SV_TAKE(r->d_qepri, double, m * sizeof(double));
    if (r->protons && r->has_qpi) SV_TAKE(r->d_qpi, double, m * sizeof(double));
    if (r->has_mass) SV_TAKE(r->d_mass, double, (size_t)r->ntr * sizeof(double));
    SV_TAKE(r->d_bg, double, (size_t)(r->nsnap + 1) * CK_NBG * r->ntr * sizeof(double));
    SV_TAKE(r->d_nsub, int, (size_t)r->nsnap * r->ntr * sizeof(int));
    SV_TAKE(r->d_non, int, (size_t)r->nsnap * r->ntr * sizeof(int));
    SV_TAKE(r->d_ric, double, (size_t)(r->nsnap + 1) * r->n * sizeof(double));
    SV_TAKE(r->d_zb, double, (size_t)(r->nsnap + 1) * 3 * sizeof(double));
#undef SV_TAKE
    if ((size_t)(p - (s->d_pool + r->off)) > r->bytes) {
        ch_set_error("serve: chunk layout exceeds its budget charge");
        return CROMA_HIP_EINVAL;
    }
    /* CMB bounds of every redshift (the nsub schedule is built on the
     * device below; cell_steps is counted once it is on the host) */
    r->cell_steps = 0;
    for (q = 0; q <= r->nsnap; q++) {
        r->h_zb[3 * q] = r->z[q];
        ch_ic_bounds(r->z[q], &r->h_zb[3 * q + 1], &r->h_zb[3 * q + 2]);
    }
    /* uploads (copy stream), IC rows on the device, ready event */
    {
        void *cs = s->copy_stream;
        int rc = 0;
        if (r->has_cre) rc |= ck_h2d(r->d_cre, r->h_cre, m * sizeof(double), cs);
        else rc |= ck_fill_f64(r->d_cre, 0.0, m, cs);
        if (r->protons) {
            if (r->has_crp) rc |= ck_h2d(r->d_crp, r->h_crp, m * sizeof(double), cs);
            else rc |= ck_fill_f64(r->d_crp, 0.0, m, cs);
        }
        if (r->has_qpi) rc |= ck_h2d(r->d_qpi, r->h_qpi, m * sizeof(double), cs);
        if (r->has_qepri) rc |= ck_h2d(r->d_qepri, r->h_qepri, m * sizeof(double), cs);
        if (r->has_mass) rc |= ck_h2d(r->d_mass, r->h_mass, (size_t)r->ntr * sizeof(double), cs);
        rc |= ck_h2d(r->d_bg, r->bg, (size_t)(r->nsnap + 1) * CK_NBG * r->ntr * sizeof(double), cs);
        /* nsub schedule of every snapshot on the device (d_nsub_one: the
         * host estimate's integers, tracer_fp_nsub.c:94-282, with the IC
         * rate at the end bins from the host's exact rows), then to the
         * pinned host copy the task ordering reads */
        for (q = 0; q < r->nsnap && !rc; q++) {
            const double *db = r->d_bg + (size_t)q * CK_NBG * r->ntr;
            double ends[2];
            ch_ctx_ic_ends(c, r->z[q], ends);
            rc |= ck_nsub_arrays(*ch_ctx_dgrid(c), r->prm, r->ntr, r->dt[q], ends[0], ends[1], db,
                                 db + (size_t)r->ntr, db + 2 * (size_t)r->ntr, db + 3 * (size_t)r->ntr,
                                 db + 4 * (size_t)r->ntr, db + 5 * (size_t)r->ntr, r->d_nsub + (size_t)q * r->ntr,
                                 r->d_non + (size_t)q * r->ntr, cs);
        }
        rc |= ck_d2h(r->h_sched, r->d_nsub, (size_t)r->nsnap * r->ntr * sizeof(int), cs);
        rc |= ck_d2h(r->h_sched + (size_t)r->nsnap * r->ntr, r->d_non, (size_t)r->nsnap * r->ntr * sizeof(int), cs);
        rc |= ck_h2d(r->d_zb, r->h_zb, (size_t)(r->nsnap + 1) * 3 * sizeof(double), cs);
        rc |= ck_ic_rows(r->n, ch_ctx_dgrid(c)->ep, g->epm1, g->epp1, r->nsnap + 1, r->d_zb, r->d_ric, 0,
                         s->d_admit_err, cs);
        rc |= ck_event_record(r->ev_ready, cs);
        if (rc) {
            ch_set_error("serve: upload: %s", ck_last_error());
            return CROMA_HIP_EDEVICE;
        }
    }
    return 0;
}

static void sv_finish(CromaHipServer *s, SvReq *r, int state, int status)
{
    pthread_mutex_lock(&s->mu);
    r->state = state;
    r->status = status;
    r->t_done = ch_now_s();
    if (state == SV_DONE) {
        s->st.completed++;
        s->st.cell_steps += r->cell_steps;
    } else if (state == SV_CANCELLED) {
        s->st.cancelled++;
    } else {
        s->st.failed++;
    }
    if (r->bytes) sv_release(s, r->off, r->bytes);
    r->bytes = 0;
    /* inputs are no longer needed; results stay until collected */
    free(r->bg);
    r->bg = 0;
    ck_host_free(r->h_in);
    r->h_in = r->h_cre = r->h_crp = r->h_qpi = r->h_qepri = r->h_mass = 0;
    ck_host_free(r->h_sched);
    r->h_sched = 0;
    if (state != SV_DONE) {
        ck_host_free(r->h_out);
        r->h_out = 0;
    }
    pthread_cond_broadcast(&s->cv_done);
    pthread_mutex_unlock(&s->mu);
}

/* Admit queued requests in arrival order while the pool has room.  A head
 * request that does not fit blocks those behind it once it has waited for
 * 16 scheduler passes (no starvation of large chunks). */
static void sv_admit(CromaHipServer *s)
{
    for (;;) {
        SvReq *r = 0;
        long long k;
        pthread_mutex_lock(&s->mu);
        for (k = s->qhead; k < s->qtail; k++) {
            SvReq *q = s->reqs[s->queue[k % s->qcap] - 1];
            size_t off;
            if (q->state != SV_QUEUED) continue;
            if (sv_cancelled(q)) {
                q->state = SV_CANCELLED;
                q->t_done = ch_now_s();
                s->st.cancelled++;
                pthread_cond_broadcast(&s->cv_done);
                continue;
            }
            if (sv_alloc(s, q->bytes, &off) == 0) {
                q->off = off;
                q->state = SV_ACTIVE;
                q->t_admit = ch_now_s();
                r = q;
                if (k == s->qhead) s->head_age = 0;
                break;
            }
            if (k == s->qhead && ++s->head_age > 16) break;
        }
        /* drop finished entries from the queue head */
        while (s->qhead < s->qtail && s->reqs[s->queue[s->qhead % s->qcap] - 1]->state != SV_QUEUED) s->qhead++;
        pthread_mutex_unlock(&s->mu);
        if (!r) return;
        {
            const int rc = sv_prepare(s, r);
            if (rc) {
                sv_finish(s, r, SV_FAILED, rc);
                continue;
            }
        }
        r->snap_next = 0;
        r->inflight = 0;
        r->elapsed = 0.0;
        r->ready = 0;
        pthread_mutex_lock(&s->mu);
        if (s->nactive == s->capactive) {
            SvReq **na = (SvReq **)realloc(s->active, (size_t)(s->capactive * 2 + 16) * sizeof(SvReq *));
            if (!na) {
                pthread_mutex_unlock(&s->mu);
                sv_finish(s, r, SV_FAILED, CROMA_HIP_ENOMEM);
                continue;
            }
            s->active = na;
            s->capactive = s->capactive * 2 + 16;
        }
        s->active[s->nactive++] = r;
        pthread_mutex_unlock(&s->mu);
    }
}

static int sv_grow_class(SvClass *k, int nslots, int ntasks, int nl)
{
    if (nslots > k->slot_cap) {
        const int cap = nslots * 2;
        ck_host_free(k->h_slots);
        ck_free(k->d_slots);
        k->h_slots = (CkSlot *)ck_host_alloc((size_t)cap * sizeof(CkSlot));
        k->d_slots = (CkSlot *)ck_malloc((size_t)cap * sizeof(CkSlot));
        if (!k->h_slots || !k->d_slots) return -1;
        k->slot_cap = cap;
    }
    if (ntasks > k->task_cap) {
        const int cap = ntasks + ntasks / 2;
        ck_host_free(k->h_tasks);
        ck_free(k->d_tasks);
        k->h_tasks = (CkTask *)ck_host_alloc((size_t)cap * sizeof(CkTask));
        k->d_tasks = (CkTask *)ck_malloc((size_t)cap * sizeof(CkTask));
        if (!k->h_tasks || !k->d_tasks) return -1;
        k->task_cap = cap;
    }
    if (nl > k->launched_cap) {
        SvReq **nlp = (SvReq **)realloc(k->launched, (size_t)nl * 2 * sizeof(SvReq *));
        if (!nlp) return -1;
        k->launched = nlp;
        k->launched_cap = nl * 2;
    }
    return 0;
}

/* One launch of class ci over every ready active request of the class. */
static int sv_launch_class(CromaHipServer *s, int ci)
{
    SvClass *k = &s->cls[ci];
    CromaHipContext *c = s->ctx[ci >> 1];
    const int protons = ci & 1;
    CromaHipTile tile;
    int i, nslots = 0, ntasks = 0, nl = 0, q, split;
    long pos, segrows = 0, nsegoff = 0;
    if (!c) return 0;
    tile = ch_ctx_tile(c, protons);
    for (i = 0; i < s->nactive; i++) {
        SvReq *r = s->active[i];
        if (r->cls != ci || r->state != SV_ACTIVE || r->snap_next >= r->nsnap || sv_cancelled(r)) continue;
        if (!r->ready) {
            const int e = ck_event_query(r->ev_ready);
            if (e == 0) continue;
            if (e < 0) return CROMA_HIP_EDEVICE;
            r->ready = 1;
            sv_count_steps(r);
        }
        r->inflight = r->nsnap - r->snap_next < tile.chain ? r->nsnap - r->snap_next : tile.chain;
        nslots += r->inflight;
        ntasks += r->ntr;
        nl++;
    }
    if (nl == 0) return 0;
    if (sv_grow_class(k, nslots, ntasks, nl)) return CROMA_HIP_ENOMEM;
    /* two species without the eta cap: proton pass, one GEMM, electron pass */
    split = protons && tile.secondary_gemm && !(ch_ctx_prm(c)->eta_dpp_cap > 0.0);
    if (split) {
        long need_off = 0, need_rows = 0;
        for (i = 0; i < s->nactive; i++) {
            SvReq *r = s->active[i];
            if (r->cls != ci || r->state != SV_ACTIVE || r->snap_next >= r->nsnap || sv_cancelled(r) || !r->ready)
                continue;
            need_off += (long)r->inflight * r->ntr;
            for (q = 0; q < r->inflight; q++) {
                const int *ns = r->h_sched + (size_t)(r->snap_next + q) * r->ntr;
                int t;
                for (t = 0; t < r->ntr; t++) need_rows += ch_choose_segments(&r->prm, ns[t], 1);
            }
        }
        if (need_off > k->segoff_cap) {
            ck_host_free(k->h_segoff);
            ck_free(k->d_segoff);
            k->h_segoff = (int *)ck_host_alloc((size_t)need_off * 2 * sizeof(int));
            k->d_segoff = (int *)ck_malloc((size_t)need_off * 2 * sizeof(int));
            k->segoff_cap = (k->h_segoff && k->d_segoff) ? need_off * 2 : 0;
            if (!k->segoff_cap) return CROMA_HIP_ENOMEM;
        }
        if (need_rows > k->segrows_cap) {
            const long cap = need_rows + need_rows / 2;
            ck_free(k->d_segrec);
            ck_free(k->d_seginje);
            k->d_segrec = (double *)ck_malloc((size_t)cap * SV_NS[ci >> 1] * sizeof(double));
            k->d_seginje = (double *)ck_malloc((size_t)cap * SV_NS[ci >> 1] * sizeof(double));
            k->segrows_cap = (k->d_segrec && k->d_seginje) ? cap : 0;
            if (!k->segrows_cap) return CROMA_HIP_ENOMEM;
        }
    }
    if (ntasks > s->tmp_cap) {
        free(s->tmp_tasks);
        free(s->keys);
        s->tmp_tasks = (CkTask *)malloc((size_t)ntasks * sizeof(CkTask));
        s->keys = (int *)malloc((size_t)ntasks * sizeof(int));
        if (!s->tmp_tasks || !s->keys) return CROMA_HIP_ENOMEM;
        s->tmp_cap = ntasks;
    }
    /* slots and unordered tasks with their cost buckets */
    nslots = 0;
    ntasks = 0;
    nl = 0;
    memset(s->hist, 0, sizeof(s->hist));
    for (i = 0; i < s->nactive; i++) {
        SvReq *r = s->active[i];
        int first, t;
        if (r->cls != ci || r->state != SV_ACTIVE || r->snap_next >= r->nsnap || sv_cancelled(r) || !r->ready)
            continue;
        first = nslots;
        for (q = 0; q < r->inflight; q++) {
            const int sn = r->snap_next + q;
            CkSlot *S = &k->h_slots[nslots++];
            int f;
            memset(S, 0, sizeof(*S));
            S->crp = r->d_crp;
            S->cre = r->d_cre;
            S->qpi = r->d_qpi;
            S->qepri = r->d_qepri;
            S->mass = r->d_mass;
            for (f = 0; f < CK_NBG; f++) {
                S->bgc[f] = r->d_bg + ((size_t)sn * CK_NBG + f) * r->ntr;
                S->bgn[f] = r->d_bg + ((size_t)(sn + 1) * CK_NBG + f) * r->ntr;
            }
            S->ric_c = r->d_ric + (size_t)sn * r->n;
            S->ric_n = r->d_ric + (size_t)(sn + 1) * r->n;
            S->nsub = r->d_nsub + (size_t)sn * r->ntr;
            S->n_on = r->d_non + (size_t)sn * r->ntr;
            S->fqe_t = protons ? ch_ctx_d_fqe(c) : 0;
            S->jp0 = ch_ctx_jp0(c);
            S->ntr = r->ntr;
            S->allow_interp = 1;
            S->protons = protons;
            S->dt_snap = r->dt[sn];
            S->elapsed = r->elapsed;
            S->prm = r->prm;
            if (split) {
                int *off = k->h_segoff + nsegoff;
                const int *ns = r->h_sched + (size_t)sn * r->ntr;
                int t;
                for (t = 0; t < r->ntr; t++) {
                    off[t] = (int)segrows;
                    segrows += ch_choose_segments(&r->prm, ns[t], 1);
                }
                S->seg_off = k->d_segoff + nsegoff;
                S->seg_rec = k->d_segrec;
                S->seg_inje = k->d_seginje;
                nsegoff += r->ntr;
            }
            r->elapsed += r->dt[sn];
        }
        for (t = 0; t < r->ntr; t++) {
            double cost = 0.0;
            for (q = 0; q < r->inflight; q++) {
                const int ns = r->h_sched[(size_t)(r->snap_next + q) * r->ntr + t];
                cost += (double)ns * (protons ? 2 : 1) + 3.0 * ch_choose_segments(&r->prm, ns, 1);
            }
            s->tmp_tasks[ntasks].slot = first;
            s->tmp_tasks[ntasks].tracer = t;
            s->tmp_tasks[ntasks].nslot = r->inflight;
            s->keys[ntasks] = ch_cost_bucket(cost);
            s->hist[s->keys[ntasks]]++;
            ntasks++;
        }
        k->launched[nl++] = r;
    }
    /* counting sort, most expensive first */
    pos = 0;
    for (i = CH_HIST - 1; i >= 0; i--) {
        const int cnt = s->hist[i];
        s->hist[i] = (int)pos;
        pos += cnt;
    }
    for (i = 0; i < ntasks; i++) k->h_tasks[s->hist[s->keys[i]]++] = s->tmp_tasks[i];
    k->nlaunched = nl;
    {
        const CkGrid *dg = ch_ctx_dgrid(c);
        const int nblocks = ch_ctx_cu_count(c) * tile.waves_per_simd;
        for (i = 0; i < nl; i++)
            if (ck_stream_wait_event(k->stream, k->launched[i]->ev_ready)) return CROMA_HIP_EDEVICE;
        const int n = SV_NS[ci >> 1];
        int rc = ck_h2d(k->d_slots, k->h_slots, (size_t)nslots * sizeof(CkSlot), k->stream) ||
                 ck_h2d(k->d_tasks, k->h_tasks, (size_t)ntasks * sizeof(CkTask), k->stream) ||
                 (split && ck_h2d(k->d_segoff, k->h_segoff, (size_t)nsegoff * sizeof(int), k->stream)) ||
                 ck_fill_i32(k->d_ctrl, 0, 1, k->stream);
        if (!rc && !split) {
            /* electron_only launches whose chunks carry no source rows run
             * the instance without source terms (the same results) */
            int mode = protons, nosrc = !protons;
            for (i = 0; i < nl && nosrc; i++)
                if (k->launched[i]->has_qepri) nosrc = 0;
            if (nosrc) mode = CK_LAUNCH_NOSRC;
            rc = ck_launch_fused(n, tile.rows_per_lane, tile.waves_per_simd, nblocks, mode, *dg, k->d_slots,
                                 k->d_tasks, ntasks, k->d_ctrl, k->d_ctrl + 1, k->stream);
        }
        if (!rc && split) {
            rc = ck_launch_fused(n, tile.rows_per_lane, tile.waves_per_simd, nblocks, 2, *dg, k->d_slots, k->d_tasks,
                                 ntasks, k->d_ctrl, k->d_ctrl + 1, k->stream);
            if (!rc && segrows > 0)
                rc = ck_sec_contract(n, ch_ctx_d_fqe(c), ch_ctx_d_npmin(c), k->d_segrec, k->d_seginje, segrows, 0,
                                     8 * ch_ctx_cu_count(c), k->stream);
            if (!rc) rc = ck_fill_i32(k->d_ctrl, 0, 1, k->stream);
            if (!rc)
                rc = ck_launch_fused(n, tile.rows_per_lane, tile.waves_per_simd, nblocks, 3, *dg, k->d_slots,
                                     k->d_tasks, ntasks, k->d_ctrl, k->d_ctrl + 1, k->stream);
        }
        if (rc || ck_d2h(k->h_err, k->d_ctrl + 1, sizeof(int), k->stream) || ck_event_record(k->ev, k->stream)) {
            ch_set_error("serve: launch: %s", ck_last_error());
            return CROMA_HIP_EDEVICE;
        }
    }
    k->busy = 1;
    s->st.iterations++;
    return 0;
}

/* Completion of class ci's launch: advance its requests; queue the result
 * download of finished ones. */
static int sv_complete_class(CromaHipServer *s, int ci)
{
    SvClass *k = &s->cls[ci];
    const int err = *k->h_err;
    int i;
    k->busy = 0;
    if (err) {
        const int code = (err & 1) ? CROMA_HIP_EPIVOT : CROMA_HIP_ENONFINITE;
        if (ck_fill_i32(k->d_ctrl + 1, 0, 1, k->stream) || ck_stream_sync(k->stream)) return CROMA_HIP_EDEVICE;
        for (i = 0; i < k->nlaunched; i++) k->launched[i]->status = code;
    }
    for (i = 0; i < k->nlaunched; i++) {
        SvReq *r = k->launched[i];
        r->snap_next += r->inflight;
        r->inflight = 0;
        if (r->status || r->snap_next >= r->nsnap) {
            const size_t m = (size_t)r->ntr * r->n;
            sv_set_state(s, r, SV_RETIRING);
            /* the class stream orders the download after the last launch */
            if (!r->status &&
                (ck_d2h(r->h_out, r->d_cre, m * sizeof(double), k->stream) ||
                 (r->protons && ck_d2h(r->h_out + m, r->d_crp, m * sizeof(double), k->stream))))
                return CROMA_HIP_EDEVICE;
            if (ck_event_record(r->ev_done, k->stream)) return CROMA_HIP_EDEVICE;
        }
    }
    return 0;
}

static void *sv_main(void *arg)
{
    CromaHipServer *s = (CromaHipServer *)arg;
    for (;;) {
        int i, any = 0, ci;
        pthread_mutex_lock(&s->mu);
        while (!s->stop && s->qhead == s->qtail && s->nactive == 0) pthread_cond_wait(&s->cv_work, &s->mu);
        if (s->stop) {
            pthread_mutex_unlock(&s->mu);
            break;
        }
        pthread_mutex_unlock(&s->mu);

        /* completions of class launches */
        for (ci = 0; ci < SV_NCLASS; ci++) {
            SvClass *k = &s->cls[ci];
            if (!k->busy) continue;
            i = ck_event_query(k->ev);
            if (i == 1) {
                const int rc = sv_complete_class(s, ci);
                if (rc) {
                    int j;
                    for (j = 0; j < k->nlaunched; j++) k->launched[j]->status = rc;
                }
                any = 1;
            } else if (i < 0) {
                int j;
                k->busy = 0;
                for (j = 0; j < k->nlaunched; j++) {
                    k->launched[j]->status = CROMA_HIP_EDEVICE;
                    sv_set_state(s, k->launched[j], SV_RETIRING);
                }
            }
        }
        /* retire requests whose results are on the host; drop cancelled ones */
        for (i = 0; i < s->nactive;) {
            SvReq *r = s->active[i];
            int done = 0, state = SV_DONE, status = r->status;
            if (r->state == SV_RETIRING) {
                const int e = ck_event_query(r->ev_done);
                if (e != 0) {
                    done = 1;
                    if (e < 0 && !status) status = CROMA_HIP_EDEVICE;
                    if (status) state = SV_FAILED;
                }
            } else if (sv_cancelled(r) && r->inflight == 0 && r->state == SV_ACTIVE) {
                /* no launch of it is on the device: its memory may be reused
                 * once its uploads have finished */
                ck_event_sync(r->ev_ready);
                done = 1;
                state = SV_CANCELLED;
                status = CROMA_HIP_ECANCELED;
            }
            if (done) {
                sv_finish(s, r, state, status);
                pthread_mutex_lock(&s->mu);
                s->active[i] = s->active[--s->nactive];
                pthread_mutex_unlock(&s->mu);
                any = 1;
            } else {
                i++;
            }
        }
        /* admissions (their uploads overlap the running launches) */
        sv_admit(s);
        /* idle classes launch their next step */
        for (ci = 0; ci < SV_NCLASS; ci++)
            if (!s->cls[ci].busy) {
                const int rc = sv_launch_class(s, ci);
                if (rc) {
                    for (i = 0; i < s->nactive; i++)
                        if (s->active[i]->cls == ci && s->active[i]->state == SV_ACTIVE) {
                            s->active[i]->status = rc;
                            sv_set_state(s, s->active[i], SV_RETIRING);
                            ck_event_record(s->active[i]->ev_done, s->copy_stream);
                        }
                }
                if (s->cls[ci].busy) any = 1;
            }
        if (!any) {
            struct timespec ts = {0, 50000};
            nanosleep(&ts, 0);
        }
    }
    return 0;
}

/* ---- public API ------------------------------------------------------ */
CromaHipServer *croma_hip_serve_create(const CromaHipParams *prm, size_t budget, double pmin, double pmax,
                                       double pemin, double pemax)
{
    CromaHipServer *s;
    int ci;
    CromaHipParams def;
    if (!prm) {
        croma_hip_default_params(&def);
        prm = &def;
    }
    if (budget < (size_t)1 << 20) {
        ch_set_error("serve: budget below 1 MiB");
        return 0;
    }
    if (ck_device_count() <= 0) {
        ch_set_error("no HIP device visible");
        return 0;
    }
    s = (CromaHipServer *)calloc(1, sizeof(*s));
    if (!s) return 0;
    s->uprm = *prm;
    s->pmin = pmin;
    s->pmax = pmax;
    s->pemin = pemin;
    s->pemax = pemax;
    pthread_mutex_init(&s->mu, 0);
    pthread_cond_init(&s->cv_work, 0);
    pthread_cond_init(&s->cv_done, 0);
    s->pool_size = budget & ~(size_t)(SV_ALIGN - 1);
    s->d_pool = (char *)ck_malloc(s->pool_size);
    s->copy_stream = ck_stream_create();
    s->d_admit_err = (int *)ck_malloc(64 * sizeof(int));
    s->capfl = 16;
    s->fl = (SvBlk *)malloc((size_t)s->capfl * sizeof(SvBlk));
    if (!s->d_pool || !s->copy_stream || !s->d_admit_err || !s->fl) {
        ch_set_error("serve: allocation of %zu bytes failed: %s", budget, ck_last_error());
        croma_hip_serve_destroy(s);
        return 0;
    }
    s->fl[0].off = 0;
    s->fl[0].size = s->pool_size;
    s->nfl = 1;
    for (ci = 0; ci < SV_NCLASS; ci++) {
        SvClass *k = &s->cls[ci];
        k->stream = ck_stream_create();
        k->ev = ck_event_create();
        k->d_ctrl = (int *)ck_malloc(64 * sizeof(int));
        k->h_err = (int *)ck_host_alloc(64);
        if (!k->stream || !k->ev || !k->d_ctrl || !k->h_err || ck_fill_i32(k->d_ctrl, 0, 64, k->stream) ||
            ck_stream_sync(k->stream)) {
            ch_set_error("serve: class setup: %s", ck_last_error());
            croma_hip_serve_destroy(s);
            return 0;
        }
        *k->h_err = 0;
    }
    if (ck_fill_i32(s->d_admit_err, 0, 64, s->copy_stream) || ck_stream_sync(s->copy_stream)) {
        croma_hip_serve_destroy(s);
        return 0;
    }
    s->st.budget_bytes = s->pool_size;
    if (pthread_create(&s->thr, 0, sv_main, s) != 0) {
        ch_set_error("serve: thread start failed");
        croma_hip_serve_destroy(s);
        return 0;
    }
    s->thr_started = 1;
    return s;
}

void croma_hip_serve_destroy(CromaHipServer *s)
{
    int ci;
    long long i;
    if (!s) return;
    if (s->thr_started) {
        pthread_mutex_lock(&s->mu);
        s->stop = 1;
        pthread_cond_broadcast(&s->cv_work);
        pthread_mutex_unlock(&s->mu);
        pthread_join(s->thr, 0);
    }
    for (ci = 0; ci < SV_NCLASS; ci++) {
        SvClass *k = &s->cls[ci];
        if (k->stream) ck_stream_sync(k->stream);
    }
    if (s->copy_stream) ck_stream_sync(s->copy_stream);
    for (ci = 0; ci < SV_NCLASS; ci++) {
        SvClass *k = &s->cls[ci];
        ck_event_destroy(k->ev);
        ck_free(k->d_ctrl);
        ck_host_free(k->h_err);
        ck_host_free(k->h_slots);
        ck_free(k->d_slots);
        ck_host_free(k->h_tasks);
        ck_free(k->d_tasks);
        ck_host_free(k->h_segoff);
        ck_free(k->d_segoff);
        ck_free(k->d_segrec);
        ck_free(k->d_seginje);
        free(k->launched);
        if (k->stream) ck_stream_destroy(k->stream);
    }
    for (i = 0; i < s->nreq; i++) {
        sv_req_free_host(s->reqs[i]);
        free(s->reqs[i]);
    }
    free(s->reqs);
    free(s->queue);
    free(s->active);
    free(s->fl);
    free(s->tmp_tasks);
    free(s->keys);
    for (ci = 0; ci < 4; ci++) croma_hip_context_destroy(s->ctx[ci]);
    ck_free(s->d_pool);
    ck_free(s->d_admit_err);
    if (s->copy_stream) ck_stream_destroy(s->copy_stream);
    pthread_mutex_destroy(&s->mu);
    pthread_cond_destroy(&s->cv_work);
    pthread_cond_destroy(&s->cv_done);
    free(s);
}

int croma_hip_serve_submit(CromaHipServer *s, const CromaHipChunk *ch, long long *id_out)
{
    SvReq *r;
    size_t m, nbg;
    int nidx, f;
    if (!s || !ch || !id_out || ch->ntracer <= 0 || ch->nsnap <= 0 || !ch->z || !ch->dt_snap || !ch->bg)
        return CROMA_HIP_EINVAL;
    nidx = sv_nidx(ch->n);
    if (nidx < 0) return CROMA_HIP_EUNSUPPORTED;
    if (!(ch->species == CROMA_HIP_SPECIES_ELECTRON || ch->species == CROMA_HIP_SPECIES_PROTON ||
          ch->species == CROMA_HIP_SPECIES_ELECTRON_PROTON))
        return CROMA_HIP_EINVAL;
    if (ch->dpp_model < 1 || ch->dpp_model > 4) return CROMA_HIP_EINVAL;
    r = (SvReq *)calloc(1, sizeof(*r));
    if (!r) return CROMA_HIP_ENOMEM;
    r->ntr = ch->ntracer;
    r->n = ch->n;
    r->nidx = nidx;
    r->nsnap = ch->nsnap;
    r->protons = ch->species != CROMA_HIP_SPECIES_ELECTRON;
    r->cls = nidx * 2 + r->protons;
    ch_params_to_ck(&s->uprm, &r->prm);
    r->prm.dpp_model = ch->dpp_model;
    r->bytes = croma_hip_chunk_bytes(r->ntr, r->n, ch->species, r->nsnap);
    if (r->bytes > s->pool_size) {
        free(r);
        ch_set_error("serve: chunk needs %zu bytes, budget %zu", croma_hip_chunk_bytes(ch->ntracer, ch->n,
                                                                                    ch->species, ch->nsnap),
                     s->pool_size);
        return CROMA_HIP_EBUDGET;
    }
    m = (size_t)r->ntr * r->n;
    nbg = (size_t)(r->nsnap + 1) * CK_NBG * r->ntr;
    r->z = (double *)malloc((size_t)(r->nsnap + 1) * sizeof(double));
    r->dt = (double *)malloc((size_t)r->nsnap * sizeof(double));
    r->bg = (double *)malloc(nbg * sizeof(double));
    r->has_cre = ch->cre != 0;
    r->has_crp = r->protons && ch->crp != 0;
    r->has_qpi = r->protons && ch->qpi != 0;
    r->has_qepri = ch->qepri != 0;
    r->has_mass = ch->mass != 0;
    {
        const size_t nin = m * (size_t)(r->has_cre + r->has_crp + r->has_qpi + r->has_qepri) +
                           (r->has_mass ? (size_t)r->ntr : 0) + 1;
        double *p;
        r->h_in = (double *)ck_host_alloc(nin * sizeof(double));
        p = r->h_in;
        if (p && r->has_cre) { r->h_cre = p; p += m; }
        if (p && r->has_crp) { r->h_crp = p; p += m; }
        if (p && r->has_qpi) { r->h_qpi = p; p += m; }
        if (p && r->has_qepri) { r->h_qepri = p; p += m; }
        if (p && r->has_mass) r->h_mass = p;
    }
    r->h_out = (double *)ck_host_alloc((r->protons ? 2 : 1) * m * sizeof(double));
    r->h_sched = (int *)ck_host_alloc(2 * (size_t)r->nsnap * r->ntr * sizeof(int));
    r->h_zb = (double *)ck_host_alloc((size_t)(r->nsnap + 1) * 3 * sizeof(double));
    r->ev_ready = ck_event_create();
    r->ev_done = ck_event_create();
    if (!r->z || !r->dt || !r->bg || !r->h_in || !r->h_out || !r->h_sched || !r->h_zb || !r->ev_ready ||
        !r->ev_done) {
        sv_req_free_host(r);
        free(r);
        return CROMA_HIP_ENOMEM;
    }
    memcpy(r->z, ch->z, (size_t)(r->nsnap + 1) * sizeof(double));
    memcpy(r->dt, ch->dt_snap, (size_t)r->nsnap * sizeof(double));
    memcpy(r->bg, ch->bg, nbg * sizeof(double));
    if (r->has_cre) memcpy(r->h_cre, ch->cre, m * sizeof(double));
    if (r->has_crp) memcpy(r->h_crp, ch->crp, m * sizeof(double));
    if (r->has_qpi) memcpy(r->h_qpi, ch->qpi, m * sizeof(double));
    if (r->has_qepri) memcpy(r->h_qepri, ch->qepri, m * sizeof(double));
    if (r->has_mass) memcpy(r->h_mass, ch->mass, (size_t)r->ntr * sizeof(double));
    if (!r->protons) {
        /* electron_only: the proton spectrum and source must be exactly zero */
        size_t i;
        for (i = 0; i < m; i++)
            if ((ch->crp && ch->crp[i] != 0.0) || (ch->qpi && ch->qpi[i] != 0.0)) {
                sv_req_free_host(r);
                free(r);
                ch_set_error("serve: electron_only chunk with a non-zero proton spectrum or source");
                return CROMA_HIP_EINVAL;
            }
    }
    for (f = 0; f <= r->nsnap; f++)
        if (!(r->z[f] >= 0.0)) {
            sv_req_free_host(r);
            free(r);
            return CROMA_HIP_EINVAL;
        }
    pthread_mutex_lock(&s->mu);
    if (s->nreq == s->capreq) {
        const long long nc = s->capreq * 2 + 64;
        SvReq **nr = (SvReq **)realloc(s->reqs, (size_t)nc * sizeof(SvReq *));
        if (!nr) {
            pthread_mutex_unlock(&s->mu);
            sv_req_free_host(r);
            free(r);
            return CROMA_HIP_ENOMEM;
        }
        s->reqs = nr;
        s->capreq = nc;
    }
    if (s->qtail - s->qhead == s->qcap) {
        const long long nc = s->qcap * 2 + 64;
        long long *nq = (long long *)malloc((size_t)nc * sizeof(long long)), k;
        if (!nq) {
            pthread_mutex_unlock(&s->mu);
            sv_req_free_host(r);
            free(r);
            return CROMA_HIP_ENOMEM;
        }
        for (k = s->qhead; k < s->qtail; k++) nq[k - s->qhead] = s->queue[k % s->qcap];
        free(s->queue);
        s->queue = nq;
        s->qtail -= s->qhead;
        s->qhead = 0;
        s->qcap = nc;
    }
    r->id = ++s->nreq;
    s->reqs[r->id - 1] = r;
    r->state = SV_QUEUED;
    r->t_submit = ch_now_s();
    s->queue[s->qtail % s->qcap] = r->id;
    s->qtail++;
    s->st.submitted++;
    *id_out = r->id;
    pthread_cond_broadcast(&s->cv_work);
    pthread_mutex_unlock(&s->mu);
    return 0;
}

static SvReq *sv_get(CromaHipServer *s, long long id)
{
    if (!s || id < 1 || id > s->nreq) return 0;
    return s->reqs[id - 1];
}

int croma_hip_serve_poll(CromaHipServer *s, long long id)
{
    SvReq *r;
    int rc;
    if (!s) return CROMA_HIP_EINVAL;
    pthread_mutex_lock(&s->mu);
    r = sv_get(s, id);
    if (!r) rc = CROMA_HIP_EINVAL;
    else if (r->state == SV_DONE) rc = 1;
    else if (r->state == SV_FAILED || r->state == SV_CANCELLED) rc = r->status ? r->status : CROMA_HIP_ECANCELED;
    else rc = 0;
    pthread_mutex_unlock(&s->mu);
    return rc;
}

int croma_hip_serve_wait(CromaHipServer *s, long long id, double *crp, double *cre, CromaHipChunkStats *st)
{
    SvReq *r;
    int rc;
    if (!s) return CROMA_HIP_EINVAL;
    pthread_mutex_lock(&s->mu);
    r = sv_get(s, id);
    if (!r) {
        pthread_mutex_unlock(&s->mu);
        return CROMA_HIP_EINVAL;
    }
    while (!(r->state == SV_DONE || r->state == SV_FAILED || r->state == SV_CANCELLED))
        pthread_cond_wait(&s->cv_done, &s->mu);
    pthread_mutex_unlock(&s->mu);
    if (st) {
        st->submit_s = r->t_submit;
        st->admit_s = r->t_admit;
        st->done_s = r->t_done;
        st->cell_steps = r->cell_steps;
        st->status = r->status;
    }
    if (r->state != SV_DONE) return r->status ? r->status : CROMA_HIP_ECANCELED;
    rc = 0;
    pthread_mutex_lock(&s->mu);
    if (r->h_out) {
        const size_t m = (size_t)r->ntr * r->n;
        if (cre) memcpy(cre, r->h_out, m * sizeof(double));
        if (crp) {
            if (r->protons) memcpy(crp, r->h_out + m, m * sizeof(double));
            else memset(crp, 0, m * sizeof(double));
        }
        /* results are handed over once */
        ck_host_free(r->h_out);
        r->h_out = 0;
    } else {
        rc = CROMA_HIP_EINVAL;
        ch_set_error("serve: results of request %lld were already collected", id);
    }
    pthread_mutex_unlock(&s->mu);
    return rc;
}

int croma_hip_serve_cancel(CromaHipServer *s, long long id)
{
    SvReq *r;
    if (!s) return CROMA_HIP_EINVAL;
    pthread_mutex_lock(&s->mu);
    r = sv_get(s, id);
    if (!r) {
        pthread_mutex_unlock(&s->mu);
        return CROMA_HIP_EINVAL;
    }
    if (r->state == SV_QUEUED || r->state == SV_ACTIVE) __atomic_store_n(&r->cancel, 1, __ATOMIC_RELEASE);
    pthread_cond_broadcast(&s->cv_work);
    pthread_mutex_unlock(&s->mu);
    return 0;
}

int croma_hip_serve_stats(CromaHipServer *s, CromaHipServeStats *out)
{
    if (!s || !out) return CROMA_HIP_EINVAL;
    pthread_mutex_lock(&s->mu);
    *out = s->st;
    out->in_use_bytes = s->in_use;
    out->peak_bytes = s->peak;
    out->queued = (int)(s->qtail - s->qhead);
    out->active = s->nactive;
    pthread_mutex_unlock(&s->mu);
    return 0;
}
