/*
 * FP64 CPU oracle for the tree-likelihood hot path.  TEST INFRASTRUCTURE ONLY:
 * used by tests/, by __graft_entry__.smoke() as the checker and by bench.py's
 * cpu_baseline leg.  The product never links or calls this file.
 *
 * A plain-C restatement of the reference's CLV engine,
 * lib/Crux/Tree/CxLik.c (Crux 1.2.0):
 *   - orc_pt           follows CxLikPt               (CxLik.c:258-292)
 *   - orc_stripe       follows CxLikExecuteStripe[Dna](CxLik.c:294-846):
 *                      step variants Compute{L,I}/Merge{L,I}, the running
 *                      max-of-contributions scale, the divide at ntrail==0,
 *                      lnScale bookkeeping, and the root aggregation with
 *                      NaN -> -inf and pattern-weight multiply.
 *   - orc_execute      follows CxLikExecute           (CxLik.c:914-968):
 *                      stripes of the site range handed to a pool of worker
 *                      threads; each stripe runs the whole plan and
 *                      recomputes P(t) per step, as the reference does.
 * Data layout is the reference's: internal CLVs [site][comp][state], leaf
 * CLVs [site][state], one lnScale per site (CxLik.h:9-44).
 *
 * The thread pool is this file's own (the reference uses a pthread message
 * queue, CxMq.c); the policy is the same: stripes are handed out to workers
 * one at a time until all are done.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

enum { ORC_COMPUTE_L = 0, ORC_COMPUTE_I = 1, ORC_MERGE_L = 2, ORC_MERGE_I = 3 };

typedef struct {
    int variant;
    int ntrail;
    double *parentMat;
    double *parentLnScale;
    const double *childMat;
    const double *childLnScale;
    double edgeLen;
} orc_step;

typedef struct {
    int dim;
    int ncomp;
    unsigned nchars;
    unsigned stripeWidth;
    unsigned nstripes;
    const unsigned *charFreqs;
    double wNorm;
    const double *compWeight; /* weightScaled per comp */
    const double *compCmult;  /* cmult per comp */
    const double *compRmult;  /* rmult of the comp's model */
    const int *compModel;     /* model index per comp */
    const double *cube;       /* per model dim^3: cube[i][j][k] */
    const double *vals;       /* per model dim eigenvalues */
    const double *piNorm;     /* per model dim normalised frequencies */
    int nsteps;
    const orc_step *steps;
    double *rootMat;
    double *rootLnScale;
    double *siteLnL;
    double *stripeLnL;
} orc_lik;

/* P[i*n+j] = max(0, sum_k cube[i][j][k] * exp(vals[k]*v))  (CxLik.c:258-292) */
void orc_pt(int n, double *P, const double *cube, const double *vals, double v)
{
    double e[64];
    for (int k = 0; k < n; k++)
        e[k] = exp(vals[k] * v);
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < n; j++) {
            double p = 0.0;
            const double *c = &cube[(size_t)i * n * n + (size_t)j * n];
            for (int k = 0; k < n; k++)
                p += c[k] * e[k];
            P[i * n + j] = p < 0.0 ? 0.0 : p;
        }
    }
}

static void orc_stripe(const orc_lik *lik, unsigned stripe)
{
    const int dim = lik->dim, ncomp = lik->ncomp;
    const unsigned dn = (unsigned)(dim * ncomp);
    const unsigned cMin = lik->stripeWidth * stripe;
    unsigned cLim = cMin + lik->stripeWidth;
    if (cLim > lik->nchars)
        cLim = lik->nchars;
    const unsigned w = cLim - cMin;
    double *P = malloc(sizeof(double) * (size_t)ncomp * dim * dim);
    double *scale = calloc(w ? w : 1, sizeof(double));

    for (int s = 0; s < lik->nsteps; s++) {
        const orc_step *st = &lik->steps[s];
        double *pm = st->parentMat, *pls = st->parentLnScale;
        const double *cm = st->childMat, *cls = st->childLnScale;
        const int leaf = (st->variant == ORC_COMPUTE_L || st->variant == ORC_MERGE_L);
        const int merge = (st->variant == ORC_MERGE_L || st->variant == ORC_MERGE_I);

        for (int mc = 0; mc < ncomp; mc++) {
            if (lik->compWeight[mc] != 0.0) {
                int m = lik->compModel[mc];
                orc_pt(dim, &P[(size_t)mc * dim * dim], &lik->cube[(size_t)m * dim * dim * dim],
                       &lik->vals[(size_t)m * dim],
                       st->edgeLen * lik->compCmult[mc] * lik->compRmult[mc] * lik->wNorm);
            }
        }
        for (unsigned c = cMin; c < cLim; c++) {
            double sc = scale[c - cMin];
            for (int mc = 0; mc < ncomp; mc++) {
                if (lik->compWeight[mc] == 0.0)
                    continue;
                const double *Pm = &P[(size_t)mc * dim * dim];
                const double *child = leaf ? &cm[(size_t)c * dim]
                                           : &cm[(size_t)c * dn + (size_t)mc * dim];
                double *par = &pm[(size_t)c * dn + (size_t)mc * dim];
                for (int i = 0; i < dim; i++) {
                    double cl = 0.0;
                    for (int j = 0; j < dim; j++)
                        cl += Pm[i * dim + j] * child[j];
                    if (cl > sc)
                        sc = cl;
                    if (merge)
                        par[i] *= cl;
                    else
                        par[i] = cl;
                }
            }
            scale[c - cMin] = sc;
            if (merge)
                pls[c] += cls[c];
            else
                pls[c] = cls[c];
        }
        if (st->ntrail == 0) {
            for (unsigned c = cMin; c < cLim; c++) {
                double se = scale[c - cMin];
                if (se != 1.0)
                    for (unsigned i = 0; i < dn; i++)
                        pm[(size_t)c * dn + i] /= se;
                pls[c] += log(se);
            }
            memset(scale, 0, sizeof(double) * (w ? w : 1));
        }
    }

    /* Root aggregation (CxLik.c:598-632). */
    double *pn = malloc(sizeof(double) * dn);
    for (int mc = 0; mc < ncomp; mc++) {
        if (lik->compWeight[mc] != 0.0) {
            int m = lik->compModel[mc];
            for (int i = 0; i < dim; i++)
                pn[mc * dim + i] = lik->piNorm[(size_t)m * dim + i] * lik->compWeight[mc];
        }
    }
    double stripeLnL = 0.0;
    for (unsigned c = cMin; c < cLim; c++) {
        double L = 0.0;
        for (int mc = 0; mc < ncomp; mc++) {
            if (lik->compWeight[mc] == 0.0)
                continue;
            for (int i = 0; i < dim; i++)
                L += pn[mc * dim + i] * lik->rootMat[(size_t)c * dn + (size_t)mc * dim + i];
        }
        double lnL = (log(L) + lik->rootLnScale[c]) * (double)lik->charFreqs[c];
        if (isnan(lnL))
            lnL = -INFINITY;
        lik->siteLnL[c] = lnL;
        stripeLnL += lnL;
    }
    lik->stripeLnL[stripe] = stripeLnL;
    free(pn);
    free(P);
    free(scale);
}

/* ------------------------------------------------------------------------ */
/* Persistent worker pool.                                                  */

typedef struct {
    pthread_mutex_t mu;
    pthread_cond_t go;
    pthread_cond_t done;
    int nworkers;
    pthread_t *threads;
    uint64_t generation;
    const orc_lik *job;
    unsigned next_stripe;
    unsigned finished;
    int quit;
} orc_pool;

static orc_pool g_pool = {PTHREAD_MUTEX_INITIALIZER, PTHREAD_COND_INITIALIZER,
                          PTHREAD_COND_INITIALIZER, 0, NULL, 0, NULL, 0, 0, 0};

static void pool_drain(orc_pool *p)
{
    /* Called with the mutex held; runs stripes until none are left. */
    const orc_lik *job = p->job;
    while (job != NULL && p->next_stripe < job->nstripes) {
        unsigned s = p->next_stripe++;
        pthread_mutex_unlock(&p->mu);
        orc_stripe(job, s);
        pthread_mutex_lock(&p->mu);
        if (++p->finished == job->nstripes)
            pthread_cond_broadcast(&p->done);
    }
}

static void *pool_worker(void *arg)
{
    orc_pool *p = arg;
    uint64_t seen = 0;
    pthread_mutex_lock(&p->mu);
    for (;;) {
        while (!p->quit && p->generation == seen)
            pthread_cond_wait(&p->go, &p->mu);
        if (p->quit)
            break;
        seen = p->generation;
        pool_drain(p);
    }
    pthread_mutex_unlock(&p->mu);
    return NULL;
}

static void pool_resize(orc_pool *p, int nworkers)
{
    if (p->nworkers == nworkers)
        return;
    if (p->nworkers > 0) {
        pthread_mutex_lock(&p->mu);
        p->quit = 1;
        pthread_cond_broadcast(&p->go);
        pthread_mutex_unlock(&p->mu);
        for (int i = 0; i < p->nworkers; i++)
            pthread_join(p->threads[i], NULL);
        free(p->threads);
        p->threads = NULL;
        p->nworkers = 0;
        p->quit = 0;
    }
    if (nworkers > 0) {
        p->threads = calloc((size_t)nworkers, sizeof(pthread_t));
        for (int i = 0; i < nworkers; i++) {
            if (pthread_create(&p->threads[i], NULL, pool_worker, p) != 0)
                break;
            p->nworkers++;
        }
    }
}

/* Run every stripe.  nthreads <= 1 runs them serially on the caller. */
void orc_execute(const orc_lik *lik, int nthreads)
{
    if (lik->nsteps <= 0)
        return;
    if (nthreads <= 1 || lik->nstripes <= 1) {
        for (unsigned s = 0; s < lik->nstripes; s++)
            orc_stripe(lik, s);
        return;
    }
    pool_resize(&g_pool, nthreads - 1);
    pthread_mutex_lock(&g_pool.mu);
    g_pool.job = lik;
    g_pool.next_stripe = 0;
    g_pool.finished = 0;
    g_pool.generation++;
    pthread_cond_broadcast(&g_pool.go);
    pool_drain(&g_pool); /* the caller works too */
    while (g_pool.finished < lik->nstripes)
        pthread_cond_wait(&g_pool.done, &g_pool.mu);
    g_pool.job = NULL;
    pthread_mutex_unlock(&g_pool.mu);
}

void orc_shutdown(void) { pool_resize(&g_pool, 0); }

/* Stripe width policy of the threaded reference (Lik.pyx:221-240): a multiple
 * of 2 sites, grown until there are at most 8*ncpus stripes. */
unsigned orc_stripe_width(unsigned nchars, unsigned ncpus)
{
    const unsigned quantum = 2;
    unsigned width = quantum;
    unsigned nstripes = nchars / width + (nchars % width != 0);
    while (nstripes > ncpus * 8) {
        width += quantum;
        nstripes = nchars / width + (nchars % width != 0);
    }
    return width;
}
