// tetra_main: the reference's driver (src/alg/critPoints.cu:1359-1514 and
// the relation requests of src/alg/test_driver.cu) on one MI355X.
//
//   1. parse the reference's options (argparse);
//   2. load TV (.vtu or a generated Kuhn grid);
//   3. build every relation whose --build_X flag is set on the device;
//      with validation on (the default bit 0), rebuild them on the host and
//      check the device results with the reference's tolerances and exactly;
//   4. classify critical points (VV + classes) and export the totals.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <iostream>

#include "argparse.h"
#include "cpu_extraction.h"
#include "critpoints.h"
#include "cuda_extraction.h"
#include "validate.h"
#include "vtk_load.h"

namespace {

struct Clock {
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double lap() {
        const auto t1 = std::chrono::steady_clock::now();
        const double s = std::chrono::duration<double>(t1 - t0).count();
        t0 = t1;
        return s;
    }
};

void say(const runtime_arguments &args, const char *what, double s) {
    if (args.debug >= DEBUG_MIN) std::cout << "[time] " << what << ": " << s << " s\n";
}

}  // namespace

int main(int argc, char *argv[]) {
    runtime_arguments args;
    parse(argc, argv, args);
    // the builders' device->host transfer and translation intervals
    set_host_timing_report(args.debug >= DEBUG_MIN);
    Clock clk;
    std::shared_ptr<TV_Data> TV = get_TV_from_VTK(args);
    say(args, "load TV", clk.lap());
    const vtkIdType nP = TV->nPoints, nC = TV->nCells;
    bool ok = true;

    // host tables the reference always builds first (VE, VF)
    VE_Data VE;
    VF_Data VF;
    TE_Data TE;
    TF_Data TF;
    const bool need_edges = args.build_TE() || args.build_EV() || args.build_ET() ||
                            args.build_FE() || args.build_EF() || args.build_EE();
    const bool need_faces = args.build_TF() || args.build_FV() || args.build_FE() ||
                            args.build_FT() || args.build_EF() || args.build_FF();
    vtkIdType nE = 0, nF = 0;
    if (need_edges) {
        nE = make_TE_and_VE(*TV, TE, VE);
        say(args, "VE [CPU]", clk.lap());
        if (args.debug >= DEBUG_MIN) std::cout << "Edges: " << nE << "\n";
    }
    if (need_faces) {
        nF = make_TF_and_VF(*TV, TF, VF);
        say(args, "VF [CPU]", clk.lap());
        if (args.debug >= DEBUG_MIN) std::cout << "Faces: " << nF << "\n";
    }
    make_TV_for_GPU(*TV);
    if (need_edges) make_VE_for_GPU(VE, nE, nP);
    if (need_faces) make_VF_for_GPU(VF, nP, nF);
    say(args, "TV upload [GPU]", clk.lap());
    const bool check = args.validate();

    if (args.build_TE()) {
        auto d = make_TE_GPU(*TV, VE, nP, nE, nC, false);
        say(args, "TE [GPU]", clk.lap());
        if (check) ok &= check_host_vs_device_TE(TE, *d) && check_exact("TE", TE, *d);
    }
    std::unique_ptr<EV_Data> hEV;
    if (args.build_EV() || (check && args.build_EE())) hEV = elective_make_EV(VE, nP, nE);
    if (args.build_EV()) {
        auto d = make_EV_GPU(VE, nP, nE, false);
        say(args, "EV [GPU]", clk.lap());
        if (check) ok &= check_host_vs_device_EV(*hEV, *d) && check_exact("EV", *hEV, *d);
    }
    if (args.build_ET()) {
        auto d = make_ET_GPU(*TV, false);
        say(args, "ET [GPU]", clk.lap());
        if (check) ok &= check_exact("ET", *elective_make_ET(TE, nE), *d);
    }
    if (args.build_TF()) {
        auto d = make_TF_GPU(*TV, VF, nP, nF, nC, false);
        say(args, "TF [GPU]", clk.lap());
        if (check) {
            TF_Data h = TF, dd = *d;
            ok &= check_exact("TF", TF, *d) && check_host_vs_device_TF(h, dd);
        }
    }
    std::unique_ptr<FV_Data> hFV;
    if (args.build_FV() || (check && (args.build_FE() || args.build_EF())))
        hFV = elective_make_FV(VF, nF);
    if (args.build_FV()) {
        auto d = make_FV_GPU(VF, nP, nF, false);
        say(args, "FV [GPU]", clk.lap());
        if (check) {
            FV_Data h = *hFV, dd = *d;
            ok &= check_exact("FV", *hFV, *d) && check_host_vs_device_FV(h, dd);
        }
    }
    std::unique_ptr<FE_Data> hFE;
    if (check && (args.build_FE() || args.build_EF())) hFE = elective_make_FE(*hFV, VE);
    if (args.build_FE()) {
        auto d = make_FE_GPU(VF, VE, nP, nE, nF, false);
        say(args, "FE [GPU]", clk.lap());
        if (check) {
            bool same = d->size() == hFE->size();
            for (size_t f = 0; same && f < d->size(); f++)
                for (int k = 0; k < 3; k++)
                    same &= (*d)[f][k].highVert == (*hFE)[f][k].highVert &&
                            (*d)[f][k].id == (*hFE)[f][k].id;
            std::cerr << "[validate] FE: " << (same ? "identical" : "DIFFERS") << "\n";
            ok &= same;
        }
    }
    std::unique_ptr<FT_Data> hFT;
    if (check && (args.build_FT() || args.build_FF())) hFT = elective_make_FT(TF, nF);
    if (args.build_FT()) {
        auto d = make_FT_GPU(*TV, false);
        say(args, "FT [GPU]", clk.lap());
        if (check) ok &= hFT && check_exact("FT", *hFT, *d);
    }
    if (args.build_EF()) {
        auto d = make_EF_GPU(*TV, false);
        say(args, "EF [GPU]", clk.lap());
        if (check) ok &= check_exact("EF", *elective_make_EF(*hFE, nE), *d);
    }
    if (args.build_VT()) {
        auto d = make_VT_GPU(*TV, false);
        say(args, "VT [GPU]", clk.lap());
        if (check) ok &= check_exact("VT", *elective_make_VT(*TV), *d);
    }
    if (args.build_TT()) {
        auto d = make_TT_GPU(*TV, false);
        say(args, "TT [GPU]", clk.lap());
        if (check) {
            auto h = elective_make_TT(*TV);
            ok &= check_host_vs_device_TT(*h, *d) && check_exact("TT", *h, *d);
        }
    }
    if (args.build_FF()) {
        auto d = make_FF_GPU(*TV, false);
        say(args, "FF [GPU]", clk.lap());
        if (check && hFT) {
            auto h = elective_make_FF(TF, *hFT);
            ok &= check_host_vs_device_FF(*h, *d) && check_exact("FF", *h, *d);
        }
    }
    if (args.build_EE()) {
        auto d = make_EE_GPU(*TV, false);
        say(args, "EE [GPU]", clk.lap());
        if (check) {
            auto h = elective_make_EE(*hEV, nP);
            ok &= check_host_vs_device_EE(*h, *d) && check_exact("EE", *h, *d);
        }
    }
    std::unique_ptr<VV_Data> hVV;
    if (args.build_VV()) {
        const int approx = args.max_VV > 0 ? args.max_VV : get_approx_max_VV(*TV, nP, args.debug);
        (void)approx;  // rows are exact CSR; the padded width only informs callers
        auto d = make_VV_GPU(*TV, (int)nC, (int)nP, false);
        say(args, "VV [GPU]", clk.lap());
        if (check) {
            hVV = elective_make_VV(*TV, nP);
            say(args, "VV [CPU]", clk.lap());
            ok &= check_host_vs_device_VV(*hVV, *d) && check_exact_VV(*hVV, *d);
        }
    }
    if (TV->vertexAttributes) {
        unsigned int *classes = gpu_critical_points(*TV, args);
        say(args, "critical points [GPU]", clk.lap());
        if (check && !args.no_partitioning && TV->partitionIDs) {
            // the same partition loop with the host builders
            unsigned int *h = partitioned_critical_points(*TV, args, false);
            say(args, "critical points [CPU, partitions]", clk.lap());
            ok &= check_exact_classes(h, classes, nP);
            delete[] h;
        } else if (check) {
            if (!hVV) hVV = elective_make_VV(*TV, nP);
            std::vector<unsigned int> h(3 * (size_t)nP);
            cpu_critpoints(*hVV, TV->vertexAttributes, nP, h.data(), args.threadNumber);
            say(args, "critical points [CPU]", clk.lap());
            ok &= check_exact_classes(h.data(), classes, nP);
        }
        if (args.export_ == "/dev/null")
            std::cerr << "Output to /dev/null omits function call entirely\n";
        else
            export_classes(classes, *TV, args);
        delete[] classes;
    }
    free_transients_GPU();
    if (check) std::cerr << (ok ? "[validate] all checks passed\n" : "[validate] CHECKS FAILED\n");
    return ok ? EXIT_SUCCESS : EXIT_FAILURE;
}
