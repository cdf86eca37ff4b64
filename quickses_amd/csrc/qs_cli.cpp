/* quickses: command-line SES mesher, a drop-in for the reference's executable
 * (main CudaSurf.cu:875-981).
 *
 *   quickses -i input.pdb -o output.obj [-v voxel] [-l passes] [-s slice] [-h]
 *
 *   -i  input PDB file (required)
 *   -o  output OBJ file (required)
 *   -v  voxel size in Angstrom (0.5)
 *   -l  Laplacian smoothing passes (1)
 *   -s  slice size of the reference's sub-grids (300); accepted for compatibility: the whole
 *       grid is one slice here and the mesh does not depend on it
 *   -h, --help
 * Values may follow the flag as the next argument or directly ("-v0.3").  Exit status: 0 on
 * success, 2 if the PDB file cannot be read, 255 on a usage error or a failed output file,
 * 1 if meshing fails (the reason is printed).
 */
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/quickses.h"

namespace {

const char *kUsage =
    "  quickses {OPTIONS}\n\n"
    "    QuickSES, SES mesh generation using GPU\n\n"
    "  OPTIONS:\n\n"
    "      -i[input.pdb]                     Input PDB file\n"
    "      -o[output.obj]                    Output OBJ mesh file\n"
    "      -l[smooth factor]                 (1) Times to run Laplacian smoothing step.\n"
    "      -v[voxel size]                    (0.5) Voxel size in Angstrom. Defines the quality of the mesh.\n"
    "      -s[slice size]                    (300) Size of the sub-grid. Defines the quantity of GPU memory\n"
    "                                        needed.\n"
    "      -h, --help                        Display this help menu\n";

bool parse_int(const char *s, int &out)
{
    char *end = nullptr;
    long v = strtol(s, &end, 10);
    if (!*s || *end) return false;
    out = (int)v;
    return true;
}

bool parse_float(const char *s, float &out)
{
    char *end = nullptr;
    float v = strtof(s, &end);
    if (!*s || *end) return false;
    out = v;
    return true;
}

}  // namespace

int main(int argc, char **argv)
{
    std::string in_path, out_path;
    float voxel = 0.5f;
    int smooth = 1, slice = 300;
    bool have_in = false, have_out = false;
    for (int a = 1; a < argc; a++) {
        const char *arg = argv[a];
        if (!strcmp(arg, "-h") || !strcmp(arg, "--help")) {
            fprintf(stderr, "%s", kUsage);
            return 0;
        }
        if (arg[0] != '-' || !arg[1] || !strchr("iolvs", arg[1])) {
            fprintf(stderr, "Flag could not be matched: %s\n%s", arg, kUsage);
            return 255;
        }
        const char flag = arg[1];
        const char *val = arg[2] ? arg + 2 : nullptr;
        if (!val) {
            if (a + 1 >= argc) {
                fprintf(stderr, "Flag '%c' requires an argument\n%s", flag, kUsage);
                return 255;
            }
            val = argv[++a];
        }
        bool ok = true;
        switch (flag) {
        case 'i': in_path = val; have_in = true; break;
        case 'o': out_path = val; have_out = true; break;
        case 'l': ok = parse_int(val, smooth); break;
        case 's': ok = parse_int(val, slice); break;
        case 'v': ok = parse_float(val, voxel); break;
        }
        if (!ok) {
            fprintf(stderr, "Argument '%s' received invalid value type '%s'\n%s", arg, val, kUsage);
            return 255;
        }
    }
    if (!have_in || !have_out) {
        fprintf(stderr, "Usage: %s", kUsage);
        return 255;
    }
    (void)slice;

    const auto t0 = std::chrono::steady_clock::now();
    float *pos = nullptr, *rad = nullptr;
    unsigned int n = 0;
    if (qs_pdb_read(in_path.c_str(), &pos, &rad, &n)) {
        fprintf(stderr, "%s\n", qs_io_error());
        return 2;
    }
    fprintf(stderr, "Grid resolution = %g\n", voxel);
    const auto t1 = std::chrono::steady_clock::now();
    fprintf(stderr, "Time for parse %g ms\n", std::chrono::duration<double, std::milli>(t1 - t0).count());

    unsigned int nv = 0, nidx = 0;
    API_computeSES(voxel, (qs_float3 *)pos, rad, n, nullptr, &nv, nullptr, &nidx, smooth);
    qs_free(pos);
    qs_free(rad);
    if (API_lastError()) {
        fprintf(stderr, "SES computation failed: %s\n", qs_last_error_message());
        API_freeMesh();
        return 1;
    }
    const auto t2 = std::chrono::steady_clock::now();
    fprintf(stderr, "Time for computing SES %g ms (%u atoms, %u vertices, %u triangles)\n",
            std::chrono::duration<double, std::milli>(t2 - t1).count(), n, nv, nidx / 3);
    const int rc = qs_obj_write(out_path.c_str(), (const float *)API_getVertices(), nv, API_getTriangles(false), nidx / 3);
    API_freeMesh();
    if (rc) {
        fprintf(stderr, "%s\n", qs_io_error());
        return 255;
    }
    return 0;
}
