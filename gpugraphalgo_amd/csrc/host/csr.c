/*
 * csr.c -- host CSR container, Galois .gr I/O, output writers, host verify.
 *
 * .gr v1 layout (read by Graph::readFromGR, graph.h:599-682):
 *   u64 version = 1, u64 sizeEdgeTy, u64 numNodes, u64 numEdges,
 *   u64 outIdx[numNodes]   cumulative END offset of each vertex's edges,
 *   u32 outs[numEdges]     destinations,
 *   u32 pad                when numEdges is odd (graph.h:642-643),
 *   u32 edgeData[numEdges] weights (graph.h:644).
 * The reference reads the weight section whatever sizeEdgeTy says; here
 * sizeEdgeTy 4 reads it and sizeEdgeTy 0 means "no weights" (all 1).
 */
#define _GNU_SOURCE
#include <errno.h>
#include <fcntl.h>
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include "gga_host.h"

static __thread char g_err[GGA_ERRBUF];

void gga_set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

void gga_clear_error(void) { g_err[0] = 0; }

const char *gga_last_error(void) { return g_err; }

const char *gga_strerror(int rc)
{
    switch (rc) {
    case GGA_OK: return "ok";
    case GGA_EINVAL: return "invalid argument";
    case GGA_EIO: return "I/O error";
    case GGA_EFORMAT: return "malformed .gr file";
    case GGA_ERANGE: return "size out of range";
    case GGA_ENOMEM: return "out of memory";
    case GGA_EHIP: return "HIP runtime error";
    case GGA_ENOSPLIT: return "graph not split";
    case GGA_ETIMEDOUT: return "iteration or barrier limit reached";
    case GGA_EBUDGET: return "device-memory budget exceeded";
    case GGA_ENODEV: return "no HIP device";
    default: return "unknown error";
    }
}

void gga_csr_free(gga_csr *g)
{
    if (!g) return;
    free(g->row_ptr);
    free(g->dst);
    free(g->w);
    memset(g, 0, sizeof *g);
}

static int csr_alloc(gga_csr *g, uint32_t n, uint32_t m)
{
    memset(g, 0, sizeof *g);
    g->n = n;
    g->m = m;
    g->row_ptr = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)n + 1));
    g->dst = (uint32_t *)malloc(sizeof(uint32_t) * (m ? m : 1));
    g->w = (uint32_t *)malloc(sizeof(uint32_t) * (m ? m : 1));
    if (!g->row_ptr || !g->dst || !g->w) {
        gga_csr_free(g);
        gga_set_error("csr_alloc: n=%u m=%u", n, m);
        return GGA_ENOMEM;
    }
    return GGA_OK;
}

int gga_csr_from_row_end(uint32_t n, uint32_t m, const uint64_t *row_end, const uint32_t *dst,
                         const uint32_t *w, gga_csr *out)
{
    if (!out || (n && !row_end) || (m && !dst)) return GGA_EINVAL;
    if (n == UINT32_MAX) return GGA_ERANGE;
    int rc = csr_alloc(out, n, m);
    if (rc) return rc;
    out->row_ptr[0] = 0;
    uint64_t prev = 0;
    for (uint32_t v = 0; v < n; v++) {
        uint64_t e = row_end[v];
        if (e < prev || e > m) {
            gga_csr_free(out);
            gga_set_error("row_end[%u]=%llu not monotone or > m=%u", v, (unsigned long long)e, m);
            return GGA_EFORMAT;
        }
        out->row_ptr[v + 1] = (uint32_t)e;
        prev = e;
    }
    if (n && prev != m) {
        gga_csr_free(out);
        gga_set_error("row_end[n-1]=%llu != m=%u", (unsigned long long)prev, m);
        return GGA_EFORMAT;
    }
    for (uint32_t e = 0; e < m; e++) {
        if (dst[e] >= n) {
            gga_csr_free(out);
            gga_set_error("edge %u: dst %u >= n %u", e, dst[e], n);
            return GGA_EFORMAT;
        }
        out->dst[e] = dst[e];
        out->w[e] = w ? w[e] : 1u;
    }
    out->has_weights = w ? 1u : 0u;
    return GGA_OK;
}

int gga_csr_read_gr(const char *path, gga_csr *out)
{
    if (!path || !out) return GGA_EINVAL;
    memset(out, 0, sizeof *out);
    int fd = open(path, O_RDONLY);
    if (fd < 0) {
        gga_set_error("open %s: %s", path, strerror(errno));
        return GGA_EIO;
    }
    struct stat st;
    if (fstat(fd, &st) != 0) {
        gga_set_error("stat %s: %s", path, strerror(errno));
        close(fd);
        return GGA_EIO;
    }
    size_t len = (size_t)st.st_size;
    if (len < 32) {
        close(fd);
        gga_set_error("%s: %zu bytes, shorter than the 32-byte header", path, len);
        return GGA_EFORMAT;
    }
    void *map = mmap(NULL, len, PROT_READ, MAP_PRIVATE, fd, 0);
    close(fd);
    if (map == MAP_FAILED) {
        gga_set_error("mmap %s: %s", path, strerror(errno));
        return GGA_EIO;
    }
    const uint64_t *h = (const uint64_t *)map;
    uint64_t version = h[0], esz = h[1], n64 = h[2], m64 = h[3];
    int rc = GGA_OK;
    if (version != 1) {
        gga_set_error("%s: version %llu != 1", path, (unsigned long long)version);
        rc = GGA_EFORMAT;
        goto out;
    }
    if (esz != 0 && esz != 4) {
        gga_set_error("%s: sizeEdgeTy %llu not 0 or 4", path, (unsigned long long)esz);
        rc = GGA_EFORMAT;
        goto out;
    }
    if (n64 >= UINT32_MAX || m64 > UINT32_MAX) {
        gga_set_error("%s: n=%llu m=%llu beyond uint32", path, (unsigned long long)n64,
                      (unsigned long long)m64);
        rc = GGA_ERANGE;
        goto out;
    }
    {
        uint64_t need = 32 + 8 * n64 + 4 * m64 + ((m64 & 1) ? 4 : 0) + (esz == 4 ? 4 * m64 : 0);
        if (len < need) {
            gga_set_error("%s: truncated (%zu < %llu bytes)", path, len, (unsigned long long)need);
            rc = GGA_EFORMAT;
            goto out;
        }
        const uint64_t *outIdx = h + 4;
        const uint32_t *outs = (const uint32_t *)(outIdx + n64);
        const uint32_t *edata = outs + m64 + ((m64 & 1) ? 1 : 0);
        rc = gga_csr_from_row_end((uint32_t)n64, (uint32_t)m64, outIdx, outs, esz == 4 ? edata : NULL,
                                  out);
    }
out:
    munmap(map, len);
    return rc;
}

static int write_all(FILE *f, const void *p, size_t bytes)
{
    return fwrite(p, 1, bytes, f) == bytes ? 0 : -1;
}

int gga_csr_write_gr(const char *path, const gga_csr *g)
{
    if (!path || !g) return GGA_EINVAL;
    FILE *f = fopen(path, "wb");
    if (!f) {
        gga_set_error("fopen %s: %s", path, strerror(errno));
        return GGA_EIO;
    }
    uint64_t hdr[4] = {1, 4, g->n, g->m};
    int bad = write_all(f, hdr, sizeof hdr);
    uint64_t buf[4096];
    for (uint32_t v = 0; v < g->n && !bad;) {
        uint32_t k = 0;
        for (; k < 4096 && v < g->n; k++, v++) buf[k] = g->row_ptr[v + 1];
        bad |= write_all(f, buf, sizeof(uint64_t) * k);
    }
    if (!bad && g->m) bad |= write_all(f, g->dst, sizeof(uint32_t) * g->m);
    if (!bad && (g->m & 1)) {
        uint32_t pad = 0;
        bad |= write_all(f, &pad, 4);
    }
    if (!bad && g->m) bad |= write_all(f, g->w, sizeof(uint32_t) * g->m);
    if (fclose(f) != 0) bad = 1;
    if (bad) {
        gga_set_error("write %s failed", path);
        return GGA_EIO;
    }
    return GGA_OK;
}

uint32_t gga_csr_maxdeg(const gga_csr *g, uint32_t *first_index)
{
    /* graph.h:338-358: strict `<` keeps the first maximum */
    uint32_t best = 0, idx = 0;
    if (g && g->n) best = g->row_ptr[1] - g->row_ptr[0];
    for (uint32_t v = 1; g && v < g->n; v++) {
        uint32_t d = g->row_ptr[v + 1] - g->row_ptr[v];
        if (best < d) {
            best = d;
            idx = v;
        }
    }
    if (first_index) *first_index = idx;
    return best;
}

/* ------------------------------------------------------------ outputs */
int gga_write_dist(const char *path, const uint32_t *dist, uint32_t n)
{
    if (!path || (n && !dist)) return GGA_EINVAL;
    FILE *f = fopen(path, "w");
    if (!f) {
        gga_set_error("fopen %s: %s", path, strerror(errno));
        return GGA_EIO;
    }
    char buf[1 << 16];
    size_t used = 0;
    int bad = 0;
    for (uint32_t i = 0; i < n && !bad; i++) {
        if (used + 12 > sizeof buf) {
            bad |= write_all(f, buf, used);
            used = 0;
        }
        used += (size_t)sprintf(buf + used, "%u\n", dist[i]);
    }
    if (!bad && used) bad |= write_all(f, buf, used);
    if (fclose(f) != 0) bad = 1;
    return bad ? GGA_EIO : GGA_OK;
}

int gga_write_result_row(const char *path, const gga_result_row *r)
{
    if (!path || !r) return GGA_EINVAL;
    FILE *f = fopen(path, "a");
    if (!f) {
        gga_set_error("fopen %s: %s", path, strerror(errno));
        return GGA_EIO;
    }
    /* field order of writeResult, myutils.h:90-109 (splitlevel printed %u) */
    fprintf(f, "%s,%s,%s,%u,%u,%u,%u,%u,%u,%u,%u,%u,%u,%u,%f,%f,%f\n", r->exename, r->algo,
            r->graphfile, r->bucket, r->nnode, r->nedge, r->ns_nnode, r->ns_nedge, r->extranode,
            r->maxdeg, (uint32_t)r->level, r->mda, r->iterations, r->runtime_ms, r->time_ms,
            r->ktime_ms, r->sktime_ms);
    return fclose(f) == 0 ? GGA_OK : GGA_EIO;
}

int gga_write_degree_dist(const char *path, uint32_t n, const uint32_t *row_ptr)
{
    if (!path || !row_ptr) return GGA_EINVAL;
    uint32_t maxd = 0;
    for (uint32_t v = 0; v < n; v++) {
        const uint32_t d = row_ptr[v + 1] - row_ptr[v];
        if (d > maxd) maxd = d;
    }
    /* dd[deg] = vertices with that out-degree, deg = 0 .. maxdeg (graph.h:98-103) */
    uint32_t *dd = (uint32_t *)calloc((size_t)maxd + 1u, sizeof *dd);
    if (!dd) return GGA_ENOMEM;
    for (uint32_t v = 0; v < n; v++) dd[row_ptr[v + 1] - row_ptr[v]]++;
    FILE *f = fopen(path, "w");
    if (!f) {
        gga_set_error("fopen %s: %s", path, strerror(errno));
        free(dd);
        return GGA_EIO;
    }
    /* "deg,count" for every degree that occurs, ascending (graph.h:105-110) */
    for (uint64_t d = 0; d <= maxd; d++)
        if (dd[d]) fprintf(f, "%llu,%u\n", (unsigned long long)d, dd[d]);
    free(dd);
    return fclose(f) == 0 ? GGA_OK : GGA_EIO;
}

void gga_graph_name(const char *path, char *out, size_t cap)
{
    if (!out || cap == 0) return;
    out[0] = 0;
    if (!path) return;
    const char *s = strrchr(path, '/');
    s = s ? s + 1 : path;
    snprintf(out, cap, "%s", s);
}

/* ------------------------------------------------------------- verify */
int gga_verify(uint32_t n, const uint32_t *row_ptr, const uint32_t *dst, const uint32_t *w,
               uint32_t src, const uint32_t *dist, uint32_t inf, gga_verify_result *r)
{
    if (!row_ptr || !dist || !r || (row_ptr[n] && !dst)) return GGA_EINVAL;
    if (inf == 0) inf = GGA_INF;
    memset(r, 0, sizeof *r);
    unsigned char *tight = (unsigned char *)calloc(n ? n : 1, 1);
    if (!tight) return GGA_ENOMEM;
    for (uint32_t u = 0; u < n; u++) {
        uint32_t du = dist[u];
        if (du >= inf) continue;
        r->reached++;
        for (uint32_t e = row_ptr[u]; e < row_ptr[u + 1]; e++) {
            uint32_t wt = w ? w[e] : 1u, v = dst[e];
            uint32_t alt = du + wt;
            if (alt < du) continue;
            if (wt > 0 && alt < inf && alt < dist[v]) r->nerr++;   /* myutils.h:12 */
            if (alt == dist[v]) tight[v] = 1;
        }
    }
    for (uint32_t v = 0; v < n; v++)
        if (v != src && dist[v] < inf && !tight[v]) r->untight++;
    r->src_ok = (src < n && dist[src] == 0) ? 1u : 0u;
    free(tight);
    return GGA_OK;
}
