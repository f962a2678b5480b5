// Critical-point driver pieces of the reference (src/alg/critPoints.cu):
// export_classes (:340-453) and the single-GPU classification run.
#include "critpoints.h"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <vector>

#include "cuda_extraction.h"
#include "tetra_c.h"

void export_classes(unsigned int *classes, TV_Data &TV, runtime_arguments &args) {
    std::ofstream file;
    std::streambuf *buf = std::cout.rdbuf();
    if (args.export_.empty()) {
        std::cerr << "No export file; outputting classes to stdout\n";
    } else {
        file.open(args.export_);
        buf = file.rdbuf();
        std::cout << "Outputting classes to " << args.export_ << "\n";
    }
    std::ostream out(buf);
    // per numeric class; the reference's labels: "minima" <- class 2,
    // "maxima" <- class 1 (its macro names, critPoints.cu:18-21)
    long long n_class[5] = {0, 0, 0, 0, 0}, insane = 0;
    for (vtkIdType v = 0; v < TV.nPoints; v++) {
        const unsigned a = classes[3 * v], b = classes[3 * v + 1], c = classes[3 * v + 2];
        // self-consistency of counts and class (critPoints.cu:379-394)
        if ((a >= 1 && b == 0 && c != 2) || (a == 0 && b >= 1 && c != 1) ||
            (a == 1 && b == 1 && c != 3) || (a > 1 && b > 1 && c != 4))
            insane++;
        if (c <= 4) n_class[c]++;
    }
    if (insane)
        std::cerr << "Insanity detected; GPU did not agree on its own answers for " << insane
                  << " points.\n";
    out << "Total number of minima:  " << n_class[2] << "\n"
        << "Total number of maxima:  " << n_class[1] << "\n"
        << "Total number of saddles: " << n_class[4] << "\n"
        << "Total number of regular: " << n_class[3] << "\n"
        << "Total number of voids:   " << n_class[0] << "\n";
    if (args.debug == DEBUG_MAX) {
        out << "Partition 0 number of minima:  " << n_class[2] << "\n"
            << "Partition 0 number of maxima:  " << n_class[1] << "\n"
            << "Partition 0 number of saddles: " << n_class[4] << "\n"
            << "Partition 0 number of regular: " << n_class[3] << "\n"
            << "Partition 0 number of voids:   " << n_class[0] << "\n";
    }
}

unsigned int *gpu_critical_points(const TV_Data &TV, const runtime_arguments &args) {
    (void)args;
    if (!TV.vertexAttributes) {
        std::fprintf(stderr, "tetra: critical points need a scalar point array\n");
        std::exit(EXIT_FAILURE);
    }
    make_TV_for_GPU(TV);
    // the mesh made by make_TV_for_GPU is the thread's current handle; the
    // C++ layer exposes it through a VV build, then classes come from the
    // C ABI on the same handle
    extern tg_mesh *tetra_current_mesh();
    tg_mesh *m = tetra_current_mesh();
    int rc = tg_vv_critpoints(m);
    unsigned int *out = new unsigned int[3 * (size_t)TV.nPoints];
    if (rc == TG_OK) rc = tg_classes_copy(m, out);
    if (rc != TG_OK) {
        std::fprintf(stderr, "tetra: critical points failed [%d]: %s\n", rc, tg_last_error());
        std::exit(EXIT_FAILURE);
    }
    return out;
}
