/*
 * ryr_cell - command-line driver with the reference's positional arguments
 *
 *   ryr_cell <Device> <name> <a3> <cjsr0> <ncx> <Ku> <r1> <Kb> <ica> <Ku2> [--flags]
 *
 * (run.sh:9,11; cuda_Cell_Detailed.cu:236-256 argv use) and the same output
 * files: wholetest_<name> (45 columns every output step,
 * cuda_Cell_Detailed.cu:1041-1072), linescan_<name> (22 columns per x,
 * :974-997) and an empty cadist_<name> (opened at :254-256, written only under
 * a macro that is off).  The compile-time variants of the Makefile become
 * flags; defaults reproduce run.sh's `make il` build (iso2 + lqt).
 *
 * Named flags:
 *   --lattice XxYxZ    interior CRUs (default 62x26x10, the reference lattice)
 *   --cl MS            pacing cycle length (default 4000, tperid :50)
 *   --nbeat N          beats (default 7, :49)    --tail MS  extra time (default 2100, :399)
 *   --nsteps N         run exactly N steps instead of nbeat*cl+tail
 *   --dt MS            time step (default 0.025, :51)
 *   --stim-dur MS / --stim-amp A / --stim-offset MS   (defaults 1 / 80 / 100, :689-692)
 *   --seed S           Philox seed (default 1)
 *   --ikr-scale X      I_Kr multiplier, svikr = 3X when not --lqt (default 1)
 *   --lqt / --no-lqt   (default on)   --iso2 / --no-iso2 (default on)   --tapsi
 *   --perm             permeabilised cell; also ci_basal 0.3 and cjsr_basal 900 as
 *                      the reference's perm build (:87-89)
 *   --iso / --slowncxchange / --randomlcc   the other -D builds (ryr_cru.h RYR_VAR_*)
 *   --nopacing         no stimulus (:694-696)
 *   --precision f64|f32   --scheme A|B
 *   --ic reference|baseline   ICs ci0/V0 0.0944/-80 (reference) or 0.1/-86 with cjsr0 = 750
 *   --cm X / --cm-fixed       Cm per 16120 CRUs (density-preserving) or a fixed Cm
 *   --outdir DIR       where the files go (default .)
 */
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include "ryr_cru.h"

static void usage()
{
    std::fprintf(stderr,
                 "usage: ryr_cell <Device> <name> <a3> <cjsr0> <ncx> <Ku> <r1> <Kb> <ica> <Ku2> [--flags]\n");
}

int main(int argc, char **argv)
{
    if (argc < 11) { usage(); return 2; }
    ryr_params p;
    ryr_cell_desc d;
    ryr_params_default(&p);
    ryr_cell_desc_default(&d);
    int device = std::atoi(argv[1]);
    if (device < 0) device = 0;
    std::string name = argv[2];
    /* argv[3] a3: clamp voltage under vclamp only (:59) */
    d.cjsr0 = std::atof(argv[4]);
    d.ncx_scale = std::atof(argv[5]);
    d.Ku = std::atof(argv[6]);
    /* argv[7] r1: passed to Compute and lcc_markov but unused (:420, reesica.h:14) */
    d.Kb = std::atof(argv[8]);
    d.ica_scale = std::atof(argv[9]);
    d.Ku2 = std::atof(argv[10]);
    int nbeat = 7;
    double tail = 2100.0;
    long long nsteps = -1;
    std::string outdir = ".";
    for (int i = 11; i < argc; ++i) {
        std::string f = argv[i];
        auto next = [&]() -> const char * {
            if (i + 1 >= argc) { std::fprintf(stderr, "missing value for %s\n", f.c_str()); std::exit(2); }
            return argv[++i];
        };
        if (f == "--lattice") {
            if (std::sscanf(next(), "%dx%dx%d", &d.nx, &d.ny, &d.nz) != 3) { usage(); return 2; }
        } else if (f == "--cl") d.cl = std::atof(next());
        else if (f == "--nbeat") nbeat = std::atoi(next());
        else if (f == "--tail") tail = std::atof(next());
        else if (f == "--nsteps") nsteps = std::atoll(next());
        else if (f == "--dt") p.dt = std::atof(next());
        else if (f == "--stim-dur") p.stim_dur = std::atof(next());
        else if (f == "--stim-amp") p.stim_amp = std::atof(next());
        else if (f == "--stim-offset") p.stim_offset = std::atof(next());
        else if (f == "--seed") d.seed = std::strtoull(next(), nullptr, 10);
        else if (f == "--ikr-scale") d.ikr_scale = std::atof(next());
        else if (f == "--lqt") p.lqt = 1;
        else if (f == "--no-lqt") p.lqt = 0;
        else if (f == "--iso2") p.iso2 = 1;
        else if (f == "--no-iso2") p.iso2 = 0;
        else if (f == "--tapsi") p.tapsi = 1;
        else if (f == "--perm") { p.variants |= RYR_VAR_PERM; p.ci0 = 0.3; d.cjsr0 = 900.0; }
        else if (f == "--iso") p.variants |= RYR_VAR_ISO;
        else if (f == "--slowncxchange") p.variants |= RYR_VAR_SLOWNCX;
        else if (f == "--randomlcc") p.variants |= RYR_VAR_RANDOMLCC;
        else if (f == "--nopacing") p.stim_amp = 0.0;
        else if (f == "--precision") {
            std::string v = next();
            p.precision = (v == "f32") ? RYR_PREC_F32 : RYR_PREC_F64;
        } else if (f == "--scheme") {
            std::string v = next();
            p.scheme = (v == "B" || v == "b") ? RYR_SCHEME_B : RYR_SCHEME_A;
        } else if (f == "--ic") {
            std::string v = next();
            if (v == "baseline") { p.ci0 = 0.1; p.v0 = -86.0; d.cjsr0 = 750.0; }
        } else if (f == "--cm") p.cm = std::atof(next());
        else if (f == "--cm-fixed") p.cm_mode = 1;
        else if (f == "--outdir") outdir = next();
        else { std::fprintf(stderr, "unknown flag %s\n", f.c_str()); usage(); return 2; }
    }
    if (nsteps < 0) nsteps = (long long)((nbeat * d.cl + tail) / p.dt + 0.5);

    ryr_handle *h = nullptr;
    int rc = ryr_create(device, &p, &d, 1, &h);
    if (rc) { std::fprintf(stderr, "ryr_create failed (%d): %s\n", rc, ryr_last_error(nullptr)); return 1; }

    std::string base = outdir + "/";
    FILE *whole = std::fopen((base + "wholetest_" + name).c_str(), "w");
    FILE *line = std::fopen((base + "linescan_" + name).c_str(), "w");
    FILE *cadist = std::fopen((base + "cadist_" + name).c_str(), "w");
    if (!whole || !line || !cadist) { std::fprintf(stderr, "cannot open output files\n"); return 1; }
    std::printf("Initialization Complete\n");
    std::fflush(stdout);

    const long long spo = p.steps_per_output;
    std::vector<double> row(RYR_DIAG_COLS), ls((size_t)d.nx * RYR_LINE_COLS);
    long long done = 0;
    while (done < nsteps) {
        long long chunk = spo - (done % spo);
        if (chunk > nsteps - done) chunk = nsteps - done;
        bool has_row = (done % spo) == 0;
        if ((rc = ryr_step(h, chunk))) break;
        done += chunk;
        if (!has_row) continue;
        int64_t st = -1;
        if ((rc = ryr_read_diag(h, row.data(), &st))) break;
        if ((rc = ryr_read_linescan(h, 0, ls.data(), (int64_t)ls.size()))) break;
        const double *r = row.data();
        auto I = [&](int k) { return (int)r[k]; };
        std::fprintf(whole, "%f %f %f %f %f  %f %f %f %f %f  "
                            "%f %f %f %f %f  %f %f %f %f %f  "
                            "%f %f %f %f %f  %f %i %i %i %i  "
                            "%i %i %i %i %i  %i %i %f %f %f  "
                            "%i %f %f %f %f\n",
                     r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7], r[8], r[9],
                     r[10], r[11], r[12], r[13], r[14], r[15], r[16], r[17], r[18], r[19],
                     r[20], r[21], r[22], r[23], r[24], r[25], I(26), I(27), I(28), I(29),
                     I(30), I(31), I(32), I(33), I(34), I(35), I(36), r[37], r[38], r[39],
                     I(40), r[41], r[42], r[43], r[44]);
        std::fflush(whole);
        for (int x = 0; x < d.nx; ++x) {
            const double *L = ls.data() + (size_t)x * RYR_LINE_COLS;
            std::fprintf(line, "%f %f %f %i %i "
                               "%i %i %i %f %f "
                               "%f %f %f %f %f "
                               "%f %f %f %f %f "
                               "%f %f\n",
                         L[0], L[1], L[2], (int)L[3], (int)L[4], (int)L[5], (int)L[6], (int)L[7],
                         L[8], L[9], L[10], L[11], L[12], L[13], L[14], L[15], L[16], L[17],
                         L[18], L[19], L[20], L[21]);
        }
        std::fprintf(line, "\n");
        std::fflush(line);
    }
    if (rc) std::fprintf(stderr, "ryr_cell: error %d: %s\n", rc, ryr_last_error(h));
    std::fclose(whole);
    std::fclose(line);
    std::fclose(cadist);
    ryr_destroy(h);
    return rc ? 1 : 0;
}
