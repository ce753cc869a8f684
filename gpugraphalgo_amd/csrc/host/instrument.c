/*
 * instrument.c -- host side of the instrumentation: the list-scheduling
 * estimate over per-workgroup timelines (the reference's combineClocks,
 * metrics.h:70-92) and the convergence report (printConv_rtime_init / _rtime /
 * _final, convergence.h:81-140).
 */
#define _GNU_SOURCE
#include <errno.h>
#include <stdlib.h>
#include <string.h>

#include "gga_host.h"

double gga_timeline_makespan(const gga_wg_record *r, uint32_t nrec, uint32_t grid, uint32_t slots)
{
    if (!r || grid == 0 || slots == 0) return 0.0;
    uint64_t *load = (uint64_t *)calloc(slots, sizeof(uint64_t));
    if (!load) return 0.0;
    uint64_t total = 0;
    for (uint32_t base = 0; base + grid <= nrec; base += grid) {
        memset(load, 0, slots * sizeof(uint64_t));
        /* in launch order, each workgroup goes to the least-loaded slot (first on ties) */
        for (uint32_t b = 0; b < grid; b++) {
            uint32_t best = 0;
            for (uint32_t k = 1; k < slots; k++)
                if (load[k] < load[best]) best = k;
            load[best] += r[base + b].ticks;
        }
        uint64_t mx = 0;
        for (uint32_t k = 0; k < slots; k++)
            if (load[k] > mx) mx = load[k];
        total += mx;
    }
    free(load);
    return (double)total / 1e5;   /* 100 MHz ticks -> ms */
}

int gga_write_conv(const char *path, int32_t level, uint32_t maxdeg, const char *graph, const char *algo,
                   const gga_conv_row *rows, uint32_t nrows, uint32_t max_dist)
{
    if (!path || (nrows && !rows)) return GGA_EINVAL;
    FILE *f = fopen(path, "a");
    if (!f) {
        gga_set_error("%s: %s", path, strerror(errno));
        return GGA_EIO;
    }
    fprintf(f, "\n\n\n----------------------------------------------------------------------------------------\n");
    fprintf(f, "Dist_Level_%d_MaxDeg_%u_%s_%s\n", level, maxdeg, graph ? graph : "", algo ? algo : "");
    for (uint32_t i = 0; i < nrows; i++)
        fprintf(f, "Iteration:%-3u = %g(%u)\n", rows[i].step, rows[i].norm2, rows[i].diff_count);
    fprintf(f, "Total Iteration : %u     Max Dist : %u\n", nrows, max_dist);
    fprintf(f, "----------------------------------------------------------------------------------------\n");
    if (fclose(f) != 0) return GGA_EIO;
    return GGA_OK;
}
