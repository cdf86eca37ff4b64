/* File formats around the SES path: the PDB atom reader and the OBJ mesh writer of the
 * reference's command-line tool, restated.
 *
 * PDB (cpdb/cpdb.cpp:34-158, 211-275 as called from CudaSurf.cu:942-968):
 *   - the file is read in chunks of at most 127 characters per line read (a 128-byte buffer);
 *   - only records starting with "ATOM  " are atoms; HETATM and the rest are skipped;
 *   - reading stops at the first record starting with "END" (END, ENDMDL, ...);
 *   - an atom whose alternate-location flag (column 17) is neither 'A' nor ' ' is dropped;
 *     hydrogens are kept;
 *   - x, y, z are the 8-character fields at columns 31, 39, 47, converted with atof;
 *   - the radius comes from the first non-blank character of the element field (columns
 *     77-79): O 1.52, C 1.70, N 1.55, H 1.20, S 1.80, P 1.80, anything else 1.40
 *     (CudaSurf.h:64-73);
 *   - atoms keep file order (the reference walks chains in order of appearance, each in file
 *     order, which is file order).
 * OBJ (writeToObj CudaSurf.cu:329-371): "v %.3f %.3f %.3f" per vertex, one empty line, then
 * "f b a c" per triangle (a, b, c) with 1-based indices: the first two corners swapped.
 */
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/quickses.h"

namespace {

float element_radius(char e)
{
    switch (e) {
    case 'O': return 1.52f;
    case 'C': return 1.70f;
    case 'N': return 1.55f;
    case 'H': return 1.20f;
    case 'S': return 1.80f;
    case 'P': return 1.80f;
    default: return 1.40f;
    }
}

/* atof of the (at most) 8 characters at `at`, stopping at the end of the line buffer */
float coord_field(const char *line, size_t len, size_t at)
{
    char buf[9] = {0};
    if (at < len) {
        size_t k = std::min<size_t>(8, len - at);
        memcpy(buf, line + at, k);
    }
    return (float)atof(buf);
}

char g_io_err[256] = "";

}  // namespace

extern "C" {

const char *qs_io_error(void) { return g_io_err; }

void qs_free(void *p) { free(p); }

int qs_pdb_read(const char *path, float **pos, float **rad, unsigned int *n)
{
    g_io_err[0] = 0;
    if (pos) *pos = nullptr;
    if (rad) *rad = nullptr;
    if (n) *n = 0;
    FILE *f = path ? fopen(path, "r") : nullptr;
    if (!f) {
        snprintf(g_io_err, sizeof(g_io_err), "pdb file can not be read: %s", path ? path : "(null)");
        return 2;
    }
    std::vector<float> P, R;
    char line[128];
    while (fgets(line, sizeof(line), f)) {
        const size_t len = strlen(line);
        if (!strncmp(line, "ATOM  ", 6)) {
            const char alt = len > 16 ? line[16] : ' ';
            if (!(alt == 'A' || alt == ' ')) continue;
            char e = 0;
            for (size_t k = 76; k <= 78 && k < len; k++)
                if (line[k] != ' ' && line[k] != '\n' && line[k] != '\r') {
                    e = line[k];
                    break;
                }
            P.push_back(coord_field(line, len, 30));
            P.push_back(coord_field(line, len, 38));
            P.push_back(coord_field(line, len, 46));
            R.push_back(element_radius(e));
        } else if (!strncmp(line, "END", 3)) {
            break;
        }
    }
    fclose(f);
    const size_t na = R.size();
    float *pp = (float *)malloc(std::max<size_t>(na, 1) * 12);
    float *rr = (float *)malloc(std::max<size_t>(na, 1) * 4);
    if (!pp || !rr) {
        free(pp);
        free(rr);
        snprintf(g_io_err, sizeof(g_io_err), "out of host memory");
        return 1;
    }
    if (na) {
        memcpy(pp, P.data(), na * 12);
        memcpy(rr, R.data(), na * 4);
    }
    if (pos) *pos = pp; else free(pp);
    if (rad) *rad = rr; else free(rr);
    if (n) *n = (unsigned)na;
    return 0;
}

int qs_obj_write(const char *path, const float *vert, unsigned int nvert, const int *tri, unsigned int ntri)
{
    g_io_err[0] = 0;
    FILE *f = path ? fopen(path, "w") : nullptr;
    if (!f) {
        snprintf(g_io_err, sizeof(g_io_err), "Failed to open output file: %s", path ? path : "(null)");
        return 2;
    }
    std::vector<char> buf(1 << 22);
    setvbuf(f, buf.data(), _IOFBF, buf.size());
    for (unsigned i = 0; i < nvert; i++)
        fprintf(f, "v %.3f %.3f %.3f\n", vert[3 * (size_t)i], vert[3 * (size_t)i + 1], vert[3 * (size_t)i + 2]);
    fprintf(f, "\n");
    for (unsigned t = 0; t < ntri; t++)
        fprintf(f, "f %d %d %d\n", tri[3 * (size_t)t + 1] + 1, tri[3 * (size_t)t] + 1, tri[3 * (size_t)t + 2] + 1);
    const int bad = ferror(f);
    if (fclose(f) != 0 || bad) {
        snprintf(g_io_err, sizeof(g_io_err), "write to %s failed", path);
        return 2;
    }
    return 0;
}

}  // extern "C"
