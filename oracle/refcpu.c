/*
 * refcpu.c -- CPU oracle for the Barnes-Hut force / time-step path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this file; only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py use it, and only as the checker (or the timed CPU baseline).
 *
 * It restates, in plain C with contraction off (-ffp-contract=off), the
 * reference's single-domain FP64 algorithm:
 *   - pointer-octree insertion        tree.f90:39-100 (BuildLocalTree),
 *                                     tree.f90:103-274 / 277-447 (Insert_Tree_e/_i)
 *   - child-point moments              tree.f90:450-553 / 556-659 (PSeudoE/I)
 *   - per-target recursive walk        tree.f90:662-717 (LocalCoulombSum),
 *                                     tree.f90:721-817 (_ee), 820-922 (_ie),
 *                                     925-1027 (_ei), 1030-1133 (_ii)
 *   - velocity Verlet halves          vverlet.f90:9-44, 49-112
 * and adds an FP64 O(N^2) direct sum with the same pair rules (like species:
 * Coulomb with r^2 > 0.01 um^2, tree.f90:806/1115; unlike species: simplified
 * Kelbg without cutoff, tree.f90:909-912).
 *
 * Every floating-point expression keeps the reference's evaluation order
 * (Fortran left-to-right association, x**2 = x*x, x**3 = (x*x)*x).
 *
 * Pinning (tests/test_oracle.py): the node tables and forces are checked
 * against the reference's own BINTREE compiled from its sources (oracle/_ref,
 * committed fixtures under tests/golden/), and against closed forms
 * (theta -> 0 equals the direct sum, 1- and 2-particle trees, centroids,
 * D == 0, finite-difference forces of the pair potentials).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_EPS 0.0014399835661145824 /* data.f90:40, eV*um */

typedef struct onode {
    int ileaf;   /* 0 twig, -1 electron leaf, +1 ion leaf (tree.f90:8) */
    int pt;      /* particle index (0-based) held by a leaf, else -1 */
    int depth;
    int count;   /* particles below (|qeff| for a twig) */
    double xx[3], lx[3], qeff, dx, D[3], Q[3], QQ[3];
    struct onode *child[8];
} onode;

typedef struct otree {
    onode *root;
    int sign;      /* -1 electrons, +1 ions */
    int nnodes;
    int error;     /* nonzero: coincident particles (unbounded recursion in the reference) */
} otree;

/* child offsets, tree.f90:109-112 (bias 1..8 -> index 0..7) */
static const double IJK[8][3] = {
    {-.25, -.25, -.25}, {+.25, -.25, -.25}, {-.25, +.25, -.25}, {+.25, +.25, -.25},
    {-.25, -.25, +.25}, {+.25, -.25, +.25}, {-.25, +.25, +.25}, {+.25, +.25, +.25}};

#define MAX_DEPTH 200

/* octant choice, tree.f90:132-173: strict '<' goes low, ties go high */
static int octant(const double *x, const onode *P)
{
    int b;
    if (x[2] < P->xx[2]) {
        if (x[1] < P->xx[1]) b = (x[0] < P->xx[0]) ? 0 : 1;
        else b = (x[0] < P->xx[0]) ? 2 : 3;
    } else {
        if (x[1] < P->xx[1]) b = (x[0] < P->xx[0]) ? 4 : 5;
        else b = (x[0] < P->xx[0]) ? 6 : 7;
    }
    return b;
}

/* Insert_Tree_e/_i, tree.f90:103-274 */
static void insert(otree *T, onode **PP, onode *R, const double *pos, int n, int bias)
{
    if (T->error) return;
    onode *P = *PP;
    if (P == NULL) {
        P = (onode *)calloc(1, sizeof(onode));
        *PP = P;
        T->nnodes++;
        P->pt = n;
        P->ileaf = T->sign;
        P->qeff = (double)T->sign;
        P->dx = 0.0;
        P->count = 1;
        P->depth = R->depth + 1;
        for (int k = 0; k < 3; k++) P->xx[k] = R->xx[k] + IJK[bias][k] * R->lx[k];
        for (int k = 0; k < 3; k++) P->lx[k] = R->lx[k] * 0.5;
        return;
    }
    if (P->depth > MAX_DEPTH) { T->error = 1; return; }
    if (P->ileaf == 0) {
        P->qeff = P->qeff + (double)T->sign;
        P->count++;
        int b = octant(pos + 3 * n, P);
        insert(T, &P->child[b], P, pos, n, b);
    } else {
        int m = P->pt;
        P->ileaf = 0;
        P->pt = -1;
        P->qeff = P->qeff + (double)T->sign;
        P->count++;
        int bm = octant(pos + 3 * m, P);
        insert(T, &P->child[bm], P, pos, m, bm);
        int bn = octant(pos + 3 * n, P);
        insert(T, &P->child[bn], P, pos, n, bn);
    }
}

/* PSeudoE/PSeudoI, tree.f90:450-553: child-point moments about the centroid */
static void pseudo(onode *P, const double *pos)
{
    if (P->ileaf != 0) {
        for (int k = 0; k < 3; k++) {
            P->xx[k] = pos[3 * P->pt + k];
            P->D[k] = 0.0; P->Q[k] = 0.0; P->QQ[k] = 0.0;
        }
        return;
    }
    double qsum = 0.0, cx[8][3], q[8], ox[3];
    int tag[8];
    for (int k = 0; k < 3; k++) { P->D[k] = 0.0; P->Q[k] = 0.0; P->QQ[k] = 0.0; }
    for (int k = 0; k < 3; k++) { ox[k] = P->xx[k]; P->xx[k] = 0.0; }
    for (int c = 0; c < 8; c++) {
        tag[c] = 0;
        onode *C = P->child[c];
        if (!C) continue;
        pseudo(C, pos);
        qsum = qsum + C->qeff;
        for (int k = 0; k < 3; k++) {
            cx[c][k] = C->xx[k];
            P->xx[k] = P->xx[k] + C->qeff * C->xx[k];
        }
        tag[c] = 1;
        q[c] = C->qeff;
    }
    for (int k = 0; k < 3; k++) P->xx[k] = P->xx[k] / qsum;
    {
        double a = P->xx[0] - ox[0], b = P->xx[1] - ox[1], d = P->xx[2] - ox[2];
        P->dx = sqrt(a * a + b * b + d * d);
    }
    for (int c = 0; c < 8; c++) {
        if (!tag[c]) continue;
        double d0 = cx[c][0] - P->xx[0], d1 = cx[c][1] - P->xx[1], d2 = cx[c][2] - P->xx[2];
        P->D[0] = P->D[0] + q[c] * d0;
        P->D[1] = P->D[1] + q[c] * d1;
        P->D[2] = P->D[2] + q[c] * d2;
        P->Q[0] = P->Q[0] + q[c] * (d0 * d0) / 2.0;
        P->Q[1] = P->Q[1] + q[c] * (d1 * d1) / 2.0;
        P->Q[2] = P->Q[2] + q[c] * (d2 * d2) / 2.0;
        P->QQ[0] = P->QQ[0] + q[c] * d0 * d1;
        P->QQ[1] = P->QQ[1] + q[c] * d1 * d2;
        P->QQ[2] = P->QQ[2] + q[c] * d2 * d0;
    }
}

static void free_nodes(onode *P)
{
    if (!P) return;
    for (int c = 0; c < 8; c++) free_nodes(P->child[c]);
    free(P);
}

/* BuildLocalTree for one species, tree.f90:46-73 */
static otree *build(int n, const double *pos, int sign)
{
    otree *T = (otree *)calloc(1, sizeof(otree));
    T->sign = sign;
    onode *R = (onode *)calloc(1, sizeof(onode));
    T->root = R;
    T->nnodes = 1;
    double mn[3], mx[3];
    for (int k = 0; k < 3; k++) { mn[k] = pos[k]; mx[k] = pos[k]; }
    for (int i = 1; i < n; i++)
        for (int k = 0; k < 3; k++) {
            double v = pos[3 * i + k];
            if (v < mn[k]) mn[k] = v;
            if (v > mx[k]) mx[k] = v;
        }
    for (int k = 0; k < 3; k++) {
        R->xx[k] = (mx[k] + mn[k]) / 2.0;
        R->lx[k] = mx[k] - mn[k];
    }
    R->ileaf = 0;
    R->dx = 0.0;
    R->qeff = 0.0; /* uninitialised in the reference (tree.f90:131); starts at 0 here -> sign*N */
    R->pt = -1;
    R->depth = 0;
    for (int i = 0; i < n && !T->error; i++) {
        R->qeff = R->qeff + (double)sign;
        R->count++;
        int b = octant(pos + 3 * i, R);
        insert(T, &R->child[b], R, pos, i, b);
    }
    if (!T->error && n > 0) pseudo(R, pos);
    return T;
}

/* ---------------------------------------------------------------- walk */

typedef struct wctx {
    const double *t;  /* target position */
    double theta, rs;
    int like;         /* 1: same species (Coulomb leaf), 0: Kelbg leaf */
    double ff_m2p;    /* target_sign * eps */
    double ff_p2p;    /* target_sign * source_sign * eps */
    double f[3];
    double epot;
    int64_t nvisit, nm2p, np2p;
    int64_t *rdf;     /* leaf-pair histogram (NULL: none, as for ee) */
    double rdx;       /* its bin width */
    int nbin;         /* Ndist + 1 bins */
} wctx;

/* RDF bin of a leaf pair, tree.f90:916-919 (INT truncates; r >= 0) */
static void rdf_count(wctx *W, double r)
{
    if (!W->rdf) return;
    int id = (int)(r / W->rdx) + 1;
    if (id <= W->nbin && id > 0) {
#pragma omp atomic
        W->rdf[id - 1]++;
    }
}

/* Search_Tree_{ee,ie,ei,ii}, tree.f90:721-1133 */
static void search(wctx *W, const onode *P)
{
    double xr[3], r2, r;
    if (P->ileaf == 0) {
        W->nvisit++;
        double s = P->lx[0];
        if (P->lx[1] > s) s = P->lx[1];
        if (P->lx[2] > s) s = P->lx[2];
        r2 = 0.0;
        for (int k = 0; k < 3; k++) {
            xr[k] = W->t[k] - P->xx[k];
            r2 = r2 + xr[k] * xr[k];
        }
        r = sqrt(r2);
        if ((s / W->theta + P->dx) < r) {
            W->nm2p++;
            double ff = W->ff_m2p;
            double R3 = r2 * r, R5 = R3 * r2, R7 = R5 * r2;
            double x = xr[0], y = xr[1], z = xr[2];
            double x2 = x * x, y2 = y * y, z2 = z * z;
            double D[6], Q[10];
            D[0] = -(1.0 / R3) + 3.0 * x2 / R5;
            D[1] = -(1.0 / R3) + 3.0 * y2 / R5;
            D[2] = -(1.0 / R3) + 3.0 * z2 / R5;
            D[3] = 3.0 * x * y / R5;
            D[4] = 3.0 * y * z / R5;
            D[5] = 3.0 * z * x / R5;
            Q[0] = 15.0 * (x2 * x) / R7 - 9.0 * x / R5;
            Q[1] = 15.0 * (y2 * y) / R7 - 9.0 * y / R5;
            Q[2] = 15.0 * (z2 * z) / R7 - 9.0 * z / R5;
            Q[3] = 15.0 * x2 * y / R7 - 3.0 * y / R5;
            Q[4] = 15.0 * x2 * z / R7 - 3.0 * z / R5;
            Q[5] = 15.0 * y2 * x / R7 - 3.0 * x / R5;
            Q[6] = 15.0 * y2 * z / R7 - 3.0 * z / R5;
            Q[7] = 15.0 * z2 * x / R7 - 3.0 * x / R5;
            Q[8] = 15.0 * z2 * y / R7 - 3.0 * y / R5;
            Q[9] = 15.0 * x * y * z / R7;
            W->epot = W->epot + 0.5 * ff * (P->qeff / r + (P->D[0] * x + P->D[1] * y + P->D[2] * z) / R3 +
                                             P->Q[0] * D[0] + P->Q[1] * D[1] + P->Q[2] * D[2] +
                                             P->QQ[0] * D[3] + P->QQ[1] * D[4] + P->QQ[2] * D[5]);
            W->f[0] = W->f[0] + ff * (P->qeff * x / R3 + D[0] * P->D[0] + D[3] * P->D[1] + D[5] * P->D[2] +
                                      Q[0] * P->Q[0] + Q[5] * P->Q[1] + Q[7] * P->Q[2] +
                                      Q[3] * P->QQ[0] + Q[4] * P->QQ[2] + Q[9] * P->QQ[1]);
            W->f[1] = W->f[1] + ff * (P->qeff * y / R3 + D[3] * P->D[0] + D[1] * P->D[1] + D[4] * P->D[2] +
                                      Q[3] * P->Q[0] + Q[1] * P->Q[1] + Q[8] * P->Q[2] +
                                      Q[5] * P->QQ[0] + Q[9] * P->QQ[2] + Q[6] * P->QQ[1]);
            W->f[2] = W->f[2] + ff * (P->qeff * z / R3 + D[5] * P->D[0] + D[4] * P->D[1] + D[2] * P->D[2] +
                                      Q[4] * P->Q[0] + Q[6] * P->Q[1] + Q[2] * P->Q[2] +
                                      Q[9] * P->QQ[0] + Q[7] * P->QQ[2] + Q[8] * P->QQ[1]);
        } else {
            for (int c = 0; c < 8; c++)
                if (P->child[c]) search(W, P->child[c]);
        }
    } else {
        r2 = 0.0;
        for (int k = 0; k < 3; k++) {
            xr[k] = W->t[k] - P->xx[k];
            r2 = r2 + xr[k] * xr[k];
        }
        if (W->like) {
            if (r2 > 0.01) {
                W->np2p++;
                r = sqrt(r2);
                double ff = W->ff_p2p;
                W->epot = W->epot + 0.5 * ff / r;
                for (int k = 0; k < 3; k++) W->f[k] = W->f[k] + ff * xr[k] / r / r2;
                rdf_count(W, r);  /* ii only: ee walks carry no histogram (tree.f90:1125-1129) */
            }
        } else {
            W->np2p++;
            r = sqrt(r2);
            double ff = W->ff_p2p;
            double e = exp(-r / W->rs);
            W->epot = W->epot + 0.5 * ff * (1.0 - e) / r;
            for (int k = 0; k < 3; k++)
                W->f[k] = W->f[k] + ff * xr[k] * (1.0 / r - e * (1.0 / r + 1.0 / W->rs)) / r2;
            rdf_count(W, r);  /* ie and ei, tree.f90:916-919, 1021-1024 */
        }
    }
}

/*
 * One force evaluation, LocalCoulombSum (tree.f90:662-717), on freshly built
 * trees (BuildLocalTree, tree.f90:39-100).  Positions are AoS [n][3].
 * Each target accumulates ee then ei (electrons) / ie then ii (ions), the
 * order in which the reference adds into ff.  epot[0] = Epot_e
 * (ee + ie + ei halves), epot[1] = Epot_i (ii).
 * counts (optional, 12 int64): {visits, m2p, p2p} for ee, ie, ei, ii.
 * Returns 0, or 1 if a tree could not be built (coincident particles).
 */
int orc_local_coulomb_sum_rdf(int ne, const double *xe, int ni, const double *xi, double theta, double rs,
                              double *fe, double *fi, double *epot, int64_t *counts, int nthreads,
                              double dx_ei, double dx_ii, int nbin, int64_t *rdf_ei, int64_t *rdf_ii);

int orc_local_coulomb_sum(int ne, const double *xe, int ni, const double *xi, double theta, double rs,
                          double *fe, double *fi, double *epot, int64_t *counts, int nthreads)
{
    return orc_local_coulomb_sum_rdf(ne, xe, ni, xi, theta, rs, fe, fi, epot, counts, nthreads, 0.0, 0.0, 0,
                                     NULL, NULL);
}

/*
 * The same, also adding every leaf pair into the RDF histograms as the
 * reference does (tree.f90:916-919, 1021-1024, 1125-1129): unlike pairs of
 * both walks (ie, ei) into rdf_ei with bin width dx_ei, like ion pairs beyond
 * the cut into rdf_ii with dx_ii; nbin = Ndist + 1 bins each (NULL: skipped).
 */
int orc_local_coulomb_sum_rdf(int ne, const double *xe, int ni, const double *xi, double theta, double rs,
                              double *fe, double *fi, double *epot, int64_t *counts, int nthreads,
                              double dx_ei, double dx_ii, int nbin, int64_t *rdf_ei, int64_t *rdf_ii)
{
    otree *Te = build(ne, xe, -1);
    otree *Ti = build(ni, xi, +1);
    int err = Te->error || Ti->error;
    double Ee = 0.0, Ei = 0.0;
    int64_t c[12] = {0};
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    if (!err) {
        /* electrons: ee (ff=-eps, Coulomb +eps) then ei (ff=-eps, Kelbg -eps) */
        double E_ee = 0.0, E_ei = 0.0, E_ie = 0.0, E_ii = 0.0;
        int64_t c0 = 0, c1 = 0, c2 = 0, c6 = 0, c7 = 0, c8 = 0;
#pragma omp parallel for schedule(static) reduction(+ : E_ee, E_ei, c0, c1, c2, c6, c7, c8)
        for (int i = 0; i < ne; i++) {
            /* Epot runs on across targets exactly as the reference's
               per-thread reduction copy does (tree.f90:683-686) */
            wctx W = {0};
            W.t = xe + 3 * i; W.theta = theta; W.rs = rs;
            W.like = 1; W.ff_m2p = -ORC_EPS; W.ff_p2p = ORC_EPS;
            W.epot = E_ee;
            search(&W, Te->root);
            E_ee = W.epot; c0 += W.nvisit; c1 += W.nm2p; c2 += W.np2p;
            W.epot = E_ei; W.nvisit = W.nm2p = W.np2p = 0;
            W.like = 0; W.ff_m2p = -ORC_EPS; W.ff_p2p = -ORC_EPS;
            W.rdf = rdf_ei; W.rdx = dx_ei; W.nbin = nbin;
            search(&W, Ti->root);
            E_ei = W.epot; c6 += W.nvisit; c7 += W.nm2p; c8 += W.np2p;
            for (int k = 0; k < 3; k++) fe[3 * i + k] = W.f[k];
        }
        int64_t c3 = 0, c4 = 0, c5 = 0, c9 = 0, c10 = 0, c11 = 0;
#pragma omp parallel for schedule(static) reduction(+ : E_ie, E_ii, c3, c4, c5, c9, c10, c11)
        for (int i = 0; i < ni; i++) {
            wctx W = {0};
            W.t = xi + 3 * i; W.theta = theta; W.rs = rs;
            W.like = 0; W.ff_m2p = ORC_EPS; W.ff_p2p = -ORC_EPS;
            W.rdf = rdf_ei; W.rdx = dx_ei; W.nbin = nbin;
            W.epot = E_ie;
            search(&W, Te->root);
            E_ie = W.epot; c3 += W.nvisit; c4 += W.nm2p; c5 += W.np2p;
            W.epot = E_ii; W.nvisit = W.nm2p = W.np2p = 0;
            W.like = 1; W.ff_m2p = ORC_EPS; W.ff_p2p = ORC_EPS;
            W.rdf = rdf_ii; W.rdx = dx_ii;
            search(&W, Ti->root);
            E_ii = W.epot; c9 += W.nvisit; c10 += W.nm2p; c11 += W.np2p;
            for (int k = 0; k < 3; k++) fi[3 * i + k] = W.f[k];
        }
        Ee = E_ee + E_ie + E_ei; /* sys%Epot_e, tree.f90:688-701 */
        Ei = E_ii;
        c[0] = c0; c[1] = c1; c[2] = c2; c[3] = c3; c[4] = c4; c[5] = c5;
        c[6] = c6; c[7] = c7; c[8] = c8; c[9] = c9; c[10] = c10; c[11] = c11;
    }
    if (epot) { epot[0] = Ee; epot[1] = Ei; }
    if (counts) memcpy(counts, c, sizeof(c));
    free_nodes(Te->root); free_nodes(Ti->root);
    free(Te); free(Ti);
    return err;
}

/* ------------------------------------------------------ tree export */

typedef struct oexport {
    int *depth, *leaf, *pt, *count, *next;
    double *xx, *lx, *qeff, *dx, *D, *Q, *QQ;
    int n, cap;
} oexport;

static int export_rec(const onode *P, oexport *E)
{
    int id = E->n++;
    if (id >= E->cap) return -1;
    E->depth[id] = P->depth;
    E->leaf[id] = P->ileaf != 0;
    E->pt[id] = P->ileaf != 0 ? P->pt : -1;
    E->count[id] = P->count;
    for (int k = 0; k < 3; k++) {
        E->xx[3 * id + k] = P->xx[k];
        E->lx[3 * id + k] = P->lx[k];
        E->D[3 * id + k] = P->D[k];
        E->Q[3 * id + k] = P->Q[k];
        E->QQ[3 * id + k] = P->QQ[k];
    }
    E->qeff[id] = P->qeff;
    E->dx[id] = P->dx;
    for (int c = 0; c < 8; c++)
        if (P->child[c] && export_rec(P->child[c], E) < 0) return -1;
    E->next[id] = E->n;
    return 0;
}

/*
 * Build one species' tree and export it in preorder (children 1..8 = the
 * reference's DFS order).  next[] is the preorder index following the node's
 * subtree.  Returns the node count, -1 if cap is too small, -2 on coincident
 * particles.
 */
int orc_export_tree(int n, const double *pos, int sign, int cap, int *depth, int *leaf, int *pt, int *count,
                    int *next, double *xx, double *lx, double *qeff, double *dx, double *D, double *Q, double *QQ)
{
    otree *T = build(n, pos, sign);
    int ret;
    if (T->error) {
        ret = -2;
    } else {
        oexport E = {depth, leaf, pt, count, next, xx, lx, qeff, dx, D, Q, QQ, 0, cap};
        ret = export_rec(T->root, &E) < 0 ? -1 : E.n;
    }
    free_nodes(T->root);
    free(T);
    return ret;
}

/* --------------------------------------------------- direct sum (FP64) */

/*
 * O(N^2) FP64 direct sum on sampled targets with the reference's pair rules.
 * tspecies[t] = 0 electron / 1 ion, tidx[t] = index in that species' array.
 * Forces out [nt][3]; pot (optional) [nt] = sum of pair energies (not halved).
 */
void orc_direct(int ne, const double *xe, int ni, const double *xi, double rs, int nt, const int *tspecies,
                const int *tidx, double *fout, double *pot, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 4)
    for (int t = 0; t < nt; t++) {
        int ts = tspecies[t];
        const double *p = ts == 0 ? xe + 3 * tidx[t] : xi + 3 * tidx[t];
        double tq = ts == 0 ? -1.0 : 1.0;
        double f[3] = {0, 0, 0}, e = 0.0;
        for (int s = 0; s < 2; s++) {
            const double *src = s == 0 ? xe : xi;
            int ns = s == 0 ? ne : ni;
            double sq = s == 0 ? -1.0 : 1.0;
            double ff = tq * sq * ORC_EPS;
            int like = (s == ts);
            for (int j = 0; j < ns; j++) {
                double xr[3], r2 = 0.0;
                for (int k = 0; k < 3; k++) {
                    xr[k] = p[k] - src[3 * j + k];
                    r2 = r2 + xr[k] * xr[k];
                }
                if (like) {
                    if (!(r2 > 0.01)) continue;
                    double r = sqrt(r2);
                    e += ff / r;
                    for (int k = 0; k < 3; k++) f[k] += ff * xr[k] / r / r2;
                } else {
                    double r = sqrt(r2);
                    double ex = exp(-r / rs);
                    e += ff * (1.0 - ex) / r;
                    for (int k = 0; k < 3; k++) f[k] += ff * xr[k] * (1.0 / r - ex * (1.0 / r + 1.0 / rs)) / r2;
                }
            }
        }
        for (int k = 0; k < 3; k++) fout[3 * t + k] = f[k];
        if (pot) pot[t] = e;
    }
}

/* ------------------------------------------------ velocity Verlet */

/* VVerletNotemp1 for one species, vverlet.f90:26-40.  x,v,f AoS [n][3]. */
void orc_vv1(int n, double *x, double *v, const double *f, double xm, double dt)
{
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) {
            v[3 * i + k] = v[3 * i + k] + 0.5 * dt * f[3 * i + k] / xm;
            x[3 * i + k] = x[3 * i + k] + dt * v[3 * i + k];
        }
}

/*
 * VVerletNotemp2 for one species, vverlet.f90:67-106.  When sample != 0 the
 * sums out[0] = sum m v^2 and out[1] = sum m (x.v)^2/|x|^2 are computed in
 * index order after the kick.
 */
void orc_vv2(int n, const double *x, double *v, const double *f, double xm, double dt, int sample, double *out)
{
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) v[3 * i + k] = v[3 * i + k] + 0.5 * dt * f[3 * i + k] / xm;
    if (sample) {
        double mv2 = 0.0, rad = 0.0;
        for (int i = 0; i < n; i++) {
            const double *xi = x + 3 * i, *vi = v + 3 * i;
            mv2 = mv2 + xm * (vi[0] * vi[0] + vi[1] * vi[1] + vi[2] * vi[2]);
            double xv = xi[0] * vi[0] + xi[1] * vi[1] + xi[2] * vi[2];
            rad = rad + xm * (xv * xv) / (xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2]);
        }
        out[0] = mv2;
        out[1] = rad;
    }
}

/*
 * Timing helper for the CPU baseline: build both trees (serial insertion as in
 * the reference) and walk a subset of targets (every `stride`-th of each
 * species) with OpenMP schedule(static).  Returns build and walk seconds.
 */
int orc_time_sample(int ne, const double *xe, int ni, const double *xi, double theta, double rs, int stride,
                    int nthreads, double *t_build, double *t_walk, int64_t *nwalked)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
    double t0 = omp_get_wtime();
#else
    double t0 = 0;
#endif
    otree *Te = build(ne, xe, -1);
    otree *Ti = build(ni, xi, +1);
#ifdef _OPENMP
    double t1 = omp_get_wtime();
#else
    double t1 = 0;
#endif
    int err = Te->error || Ti->error;
    double sink = 0.0;
    int64_t nw = 0;
    if (!err) {
#pragma omp parallel for schedule(static) reduction(+ : sink, nw)
        for (int i = 0; i < ne; i += stride) {
            wctx W = {0};
            W.t = xe + 3 * i; W.theta = theta; W.rs = rs;
            W.like = 1; W.ff_m2p = -ORC_EPS; W.ff_p2p = ORC_EPS;
            search(&W, Te->root);
            W.like = 0; W.ff_p2p = -ORC_EPS;
            search(&W, Ti->root);
            sink += W.f[0] + W.epot;
            nw++;
        }
#pragma omp parallel for schedule(static) reduction(+ : sink, nw)
        for (int i = 0; i < ni; i += stride) {
            wctx W = {0};
            W.t = xi + 3 * i; W.theta = theta; W.rs = rs;
            W.like = 0; W.ff_m2p = ORC_EPS; W.ff_p2p = -ORC_EPS;
            search(&W, Te->root);
            W.like = 1; W.ff_p2p = ORC_EPS;
            search(&W, Ti->root);
            sink += W.f[0] + W.epot;
            nw++;
        }
    }
#ifdef _OPENMP
    double t2 = omp_get_wtime();
#else
    double t2 = 0;
#endif
    *t_build = t1 - t0;
    *t_walk = t2 - t1;
    *nwalked = nw;
    free_nodes(Te->root); free_nodes(Ti->root);
    free(Te); free(Ti);
    return err ? -1 : (sink != sink ? 1 : 0);
}

/*
 * The per-target walk of orc_local_coulomb_sum for a list of targets only
 * (tspecies[t] 0 electron / 1 ion, tidx[t] index in that species): the same
 * trees and the same search order, so each listed target gets exactly the
 * force the full walk gives it.  Lets tests pin the FP64 walk at sizes where
 * walking every target on the CPU would take minutes.
 */
int orc_force_targets(int ne, const double *xe, int ni, const double *xi, double theta, double rs, int nt,
                      const int *tspecies, const int *tidx, double *f, int nthreads)
{
    otree *Te = build(ne, xe, -1);
    otree *Ti = build(ni, xi, +1);
    int err = Te->error || Ti->error;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    if (!err) {
#pragma omp parallel for schedule(static)
        for (int t = 0; t < nt; t++) {
            wctx W = {0};
            W.theta = theta; W.rs = rs;
            if (tspecies[t] == 0) {
                W.t = xe + 3 * tidx[t];
                W.like = 1; W.ff_m2p = -ORC_EPS; W.ff_p2p = ORC_EPS;
                search(&W, Te->root);
                W.like = 0; W.ff_m2p = -ORC_EPS; W.ff_p2p = -ORC_EPS;
                search(&W, Ti->root);
            } else {
                W.t = xi + 3 * tidx[t];
                W.like = 0; W.ff_m2p = ORC_EPS; W.ff_p2p = -ORC_EPS;
                search(&W, Te->root);
                W.like = 1; W.ff_m2p = ORC_EPS; W.ff_p2p = ORC_EPS;
                search(&W, Ti->root);
            }
            for (int k = 0; k < 3; k++) f[3 * t + k] = W.f[k];
        }
    }
    free_nodes(Te->root); free_nodes(Ti->root);
    free(Te); free(Ti);
    return err;
}
