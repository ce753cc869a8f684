/*
 * cli.c -- gga_run: drop-in for the reference executables
 *     {n,na,d,da}_{bfs,sssp}.exe <g.gr> [L] [bucket | outfile]   (bfs_sssp.cu:103-539)
 *
 * Positional arguments keep the reference's meaning (bfs_sssp.cu:149-185):
 *   L           NodeSplitLevel: MDA = ceil(maxdeg / L); 0 = no split;
 *               -1 = histogram heuristic with `bucket` buckets (default 10)
 *   3rd arg     bucket count when L == -1, otherwise the distance output file
 * Options (may appear anywhere):
 *   --algo bfs|sssp     default: "bfs" if the program name contains it, else sssp
 *   --mode topo|worklist|persistent|auto   (default auto)
 *   --mda T             explicit split threshold (overrides L)
 *   --src s             source vertex (the reference hard-codes 0, bfs_sssp.cu:203)
 *   --ref-inf           unreached = 1000000000 (MYINFINITY) instead of 4294967295
 *   --out FILE          distance file (same as the positional outfile)
 *   --result FILE       CSV row file (default result.vi, myutils.h:89)
 *   --device d          HIP device (default 0)
 *   --dd                also write DD_<graph>.txt, the out-degree distribution
 *                       (printfDegreeDistribution, graph.h:98-114; bfs_sssp.cu:196-200)
 * Output: the reference's stdout lines, the Ary2File distance file, and one
 * 17-field writeResult row.
 */
#define _GNU_SOURCE
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../../../include/gga.h"

static double now_ms(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

static void usage(const char *prog)
{
    printf("Usage: %s <file>\n", prog);   /* bfs_sssp.cu:150 */
}

int main(int argc, char **argv)
{
    const char *pos[3] = {NULL, NULL, NULL};
    int npos = 0;
    const char *algo = NULL, *mode_s = "auto", *outfile = NULL, *result = "result.vi";
    uint32_t mda = 0, src = 0, inf = GGA_INF;
    int device = 0, want_dd = 0;
    for (int i = 1; i < argc; i++) {
        const char *a = argv[i];
        if (!strcmp(a, "--algo") && i + 1 < argc) algo = argv[++i];
        else if (!strcmp(a, "--mode") && i + 1 < argc) mode_s = argv[++i];
        else if (!strcmp(a, "--mda") && i + 1 < argc) mda = (uint32_t)strtoul(argv[++i], NULL, 10);
        else if (!strcmp(a, "--src") && i + 1 < argc) src = (uint32_t)strtoul(argv[++i], NULL, 10);
        else if (!strcmp(a, "--ref-inf")) inf = GGA_REF_INF;
        else if (!strcmp(a, "--out") && i + 1 < argc) outfile = argv[++i];
        else if (!strcmp(a, "--result") && i + 1 < argc) result = argv[++i];
        else if (!strcmp(a, "--device") && i + 1 < argc) device = atoi(argv[++i]);
        else if (!strcmp(a, "--dd")) want_dd = 1;
        else if (npos < 3) pos[npos++] = a;
    }
    if (!algo) algo = strstr(argv[0], "bfs") ? "bfs" : "sssp";
    const int bfs = !strcmp(algo, "bfs");
    printf("%s\n", bfs ? "BFS" : "SSSP");
    fflush(stdout);
    if (npos < 1) {
        usage(argv[0]);
        return 1;
    }
    printf("dc=%d\n", gga_device_count());
    printf("Device Name:%s\n", gga_version());

    int split_level = 1;     /* argc == 2: no split (EnableNodeSplit = 0) */
    int enable_split = 0;
    uint32_t bucket = 10;
    if (npos > 1) {
        enable_split = 1;
        split_level = atoi(pos[1]);
        if (split_level == 0) enable_split = 0;
        if (split_level == -1 && npos > 2) bucket = (uint32_t)atoi(pos[2]);
    }
    if (split_level != -1 && npos > 2 && !outfile) outfile = pos[2];

    char gname[200];
    gga_graph_name(pos[0], gname, sizeof gname);
    printf("%s\n%s\n", gname, pos[0]);

    gga_graph *g = NULL;
    int rc;
    if (want_dd) {
        /* the histogram needs the host CSR: read it here, write DD_<graph>.txt
         * (before the split, as bfs_sssp.cu:196-200), then upload it */
        gga_csr c;
        rc = gga_csr_read_gr(pos[0], &c);
        if (!rc) {
            char ddname[256];
            snprintf(ddname, sizeof ddname, "DD_%s.txt", gname);
            rc = gga_write_degree_dist(ddname, c.n, c.row_ptr);
            if (!rc) rc = gga_graph_from_csr32(c.n, c.m, c.row_ptr, c.dst, c.has_weights ? c.w : NULL, device, &g);
            gga_csr_free(&c);
        }
    } else {
        rc = gga_graph_load_gr(pos[0], device, &g);
    }
    if (rc) {
        fprintf(stderr, "load %s: %s (%s)\n", pos[0], gga_strerror(rc), gga_last_error());
        return 2;
    }
    printf("nnodes=%u, nedges=%u.\n", gga_graph_n(g), gga_graph_m(g));

    gga_split_stats ss;
    if (mda) rc = gga_graph_split(g, mda, 0, 0, &ss);
    else if (enable_split) rc = gga_graph_split(g, 0, split_level, bucket, &ss);
    else rc = gga_graph_split(g, 0, 0, 0, &ss);
    if (rc) {
        fprintf(stderr, "split: %s (%s)\n", gga_strerror(rc), gga_last_error());
        gga_graph_free(g);
        return 2;
    }
    printf("MaxDeg=%u\n", ss.maxdeg);
    printf("maxOutDegree=%u\n", ss.mda);
    printf("After node split:\nnnodes=%u,nedges=%u (Extranodes=%u)\n", ss.ns_nnode, ss.ns_nedge,
           ss.extranode);

    gga_opts o;
    gga_opts_init(&o);
    o.inf = inf;
    o.count = 1;   /* the reference's ktime line needs the per-launch times */
    if (!strcmp(mode_s, "topo")) o.mode = GGA_MODE_TOPO;
    else if (!strcmp(mode_s, "worklist")) o.mode = GGA_MODE_WORKLIST;
    else if (!strcmp(mode_s, "persistent")) o.mode = GGA_MODE_PERSISTENT;
    else o.mode = GGA_MODE_AUTO;

    uint32_t n = gga_graph_n(g);
    uint32_t *dist = (uint32_t *)malloc(sizeof(uint32_t) * n);
    if (!dist) {
        gga_graph_free(g);
        return 2;
    }
    printf("solving.\n");
    gga_stats st;
    double t0 = now_ms();
    rc = bfs ? gga_bfs(g, src, dist, &o, &st) : gga_sssp(g, src, dist, &o, &st);
    double t1 = now_ms();
    if (rc) {
        fprintf(stderr, "%s: %s (%s)\n", algo, gga_strerror(rc), gga_last_error());
        free(dist);
        gga_graph_free(g);
        return 2;
    }
    printf("iterations = %u.\n", st.iterations);
    if (outfile) {
        printf("Writing Result To file : %s\n", outfile);
        fflush(stdout);
        rc = gga_write_dist(outfile, dist, n);
        if (rc) fprintf(stderr, "write %s: %s\n", outfile, gga_strerror(rc));
    }
    gga_verify_result vr;
    rc = gga_graph_verify(g, src, bfs, inf, &vr);
    if (rc) fprintf(stderr, "verify: %s (%s)\n", gga_strerror(rc), gga_last_error());
    printf("verifying: no of errors = %llu.\n", (unsigned long long)vr.nerr);
    uint32_t runtime = (uint32_t)(t1 - t0);
    printf("%u ms.\n", runtime);
    printf("Time:%.2f ktime=%.2f sktime=%.2f\n", st.time_ms, st.ktime_ms, st.sktime_ms);
    fflush(stdout);

    gga_result_row row;
    memset(&row, 0, sizeof row);
    snprintf(row.exename, sizeof row.exename, "%s", argv[0]);
    snprintf(row.algo, sizeof row.algo, "%s", bfs ? "BFS" : "SSSP");
    snprintf(row.graphfile, sizeof row.graphfile, "%s", gname);
    row.bucket = split_level < 0 ? bucket : 0;   /* bfs_sssp.cu:177-180 */
    row.nnode = ss.nnode;
    row.nedge = ss.nedge;
    row.ns_nnode = ss.ns_nnode;
    row.ns_nedge = ss.ns_nedge;
    row.extranode = ss.extranode;
    row.maxdeg = ss.maxdeg;
    row.level = split_level;
    row.mda = enable_split || mda ? ss.mda : 0;   /* unset without a split */
    row.iterations = st.iterations;
    row.runtime_ms = runtime;
    row.time_ms = st.time_ms;
    row.ktime_ms = st.ktime_ms;
    row.sktime_ms = st.sktime_ms;
    gga_write_result_row(result, &row);
    free(dist);
    gga_graph_free(g);
    return 0;
}
