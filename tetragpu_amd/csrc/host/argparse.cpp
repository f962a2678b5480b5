// Command line of the reference driver (src/argparse.cu:52-326): the same
// options, help texts, defaults and flag-bit semantics.  --validate is
// always available here; bit 0 starts at the bitset default like every other
// bit (flags can only be switched on), and --no_validate (an option of this
// library) switches validation off.  --export_vertices (also this library's)
// adds the per-vertex lines to the class export.
#include "argparse.h"

#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <sstream>

#include "tetra_c.h"

std::string usage(const char *argv0, const struct option *options, const option_map &help_info,
                  const option_map &metavars) {
    auto lookup = [](const option_map &m, const char *key) -> std::string {
        for (const auto &kv : m)
            if (std::string(kv.first) == key) return kv.second;
        return "";
    };
    std::stringstream out;
    out << "Usage: " << argv0 << " [options]\n";
    for (const struct option *o = options; o->name; o++) {
        out << "  ";
        if (!o->flag && o->val > 0 && o->val < 128) out << "-" << (char)o->val << " | ";
        out << "--" << o->name;
        const std::string meta = lookup(metavars, o->name);
        if (o->has_arg == required_argument) out << " " << (meta.empty() ? "VALUE" : meta);
        const std::string help = lookup(help_info, o->name);
        if (!help.empty()) out << "\n      " << help;
        out << "\n";
    }
    return out.str();
}

void parse(int argc, char *argv[], runtime_arguments &args) {
    int bits[14] = {0};
    int no_validate = 0;
    opterr = 0;
    optind = 1;
    const char *short_opts = "hi:t:e:a:p:l:v:g:d:n";
    static struct option long_opts[] = {
        {"help", no_argument, 0, 'h'},
        {"input", required_argument, 0, 'i'},
        {"threads", required_argument, 0, 't'},
        {"export", required_argument, 0, 'e'},
        {"arrayname", required_argument, 0, 'a'},
        {"partitioningname", required_argument, 0, 'p'},
        {"thread_logname", required_argument, 0, 'l'},
        {"max_VV", required_argument, 0, 'v'},
        {"gpus", required_argument, 0, 'g'},
        {"debug", required_argument, 0, 'd'},
        {"no_partitioning", no_argument, 0, 'n'},
        {"validate", no_argument, nullptr, 0},
        {"no_validate", no_argument, nullptr, 0},
        {"build_TE", no_argument, nullptr, 0},
        {"build_EV", no_argument, nullptr, 0},
        {"build_ET", no_argument, nullptr, 0},
        {"build_TF", no_argument, nullptr, 0},
        {"build_FV", no_argument, nullptr, 0},
        {"build_FE", no_argument, nullptr, 0},
        {"build_FT", no_argument, nullptr, 0},
        {"build_EF", no_argument, nullptr, 0},
        {"build_VT", no_argument, nullptr, 0},
        {"build_TT", no_argument, nullptr, 0},
        {"build_FF", no_argument, nullptr, 0},
        {"build_EE", no_argument, nullptr, 0},
        {"build_VV", no_argument, nullptr, 0},
        {"export_vertices", no_argument, nullptr, 0},
        {0, 0, 0, 0}};
    // flag-only options point at their bit (bit 0 validate, 1..13 relations)
    long_opts[11].flag = &bits[0];
    long_opts[11].val = 1;
    long_opts[12].flag = &no_validate;
    long_opts[12].val = 1;
    for (int b = 1; b <= 13; b++) {
        long_opts[12 + b].flag = &bits[b];
        long_opts[12 + b].val = 1;
    }
    int export_vertices = 0;
    long_opts[26].flag = &export_vertices;
    long_opts[26].val = 1;
    const option_map help_info = {
        {"help", "Print this help message and exit"},
        {"input", "Tetrahedral mesh input (.vtu, or kuhn:N[:seed[:scalar]] for a generated grid)"},
        {"threads", "CPU threads per GPU"},
        {"export", "File to export CritPoints classifications to"},
        {"arrayname", "Array to use for scalar data (as string name)"},
        {"partitioningname", "Array to use for multi-GPU partitioning (as string name)"},
        {"thread_logname", "Prefix for per-thread logs (suffixed as 'thread_#.log'), leave "
                           "undefined to allow all threads to write to std::cout"},
        {"max_VV", "Override estimation of max VV with integer value"},
        {"gpus", "Set number of GPUs to use (this build runs on one GPU)"},
        {"debug", "Set the debug level (0=OFF, 1=Reduced, 2=Verbose)"},
        {"no_partitioning", "Disable partitioning; requires all memory to fit within single GPU"},
        {"validate", "Check GPU results using CPU"},
        {"no_validate", "Skip the CPU check (validation is on by default)"},
        {"build_TE", "Build the TE relationship"}, {"build_EV", "Build the EV relationship"},
        {"build_ET", "Build the ET relationship"}, {"build_TF", "Build the TF relationship"},
        {"build_FV", "Build the FV relationship"}, {"build_FE", "Build the FE relationship"},
        {"build_FT", "Build the FT relationship"}, {"build_EF", "Build the EF relationship"},
        {"build_VT", "Build the VT relationship"}, {"build_TT", "Build the TT relationship"},
        {"build_FF", "Build the FF relationship"}, {"build_EE", "Build the EE relationship"},
        {"build_VV", "Build the VV relationship"},
        {"export_vertices", "Also export one 'A <vertex> <class>' line per classified vertex"},
    };
    const option_map metavars = {
        {"input", "input.vtu"},
        {"export", "classes.txt"},
        {"arrayname", "my_scalar_data_name"},
        {"partitioningname", "my_partition_data_name"},
        {"thread_logname", "my_log_name"},
        {"max_VV", "(INT>0, preferably multiple of 32)"},
        {"gpus", "(INT>=0)"},
        {"debug", "{0,1,2}"},
    };
    std::stringstream errors;
    int bad = 0;
    while (true) {
        int idx = 0;
        const int c = getopt_long(argc, argv, short_opts, long_opts, &idx);
        if (c == -1) break;
        switch (c) {
        case 0:
            break;
        case '?':
            errors << "Unrecognized argument: " << argv[optind - 1] << "\n";
            bad++;
            break;
        case 'i':
            args.fileName = optarg;
            break;
        case 't':
            args.threadNumber = std::atoi(optarg);
            if (args.threadNumber <= 0) {
                std::cerr << "Thread argument must be integer >= 1\n";
                bad++;
            }
            break;
        case 'e':
            args.export_ = optarg;
            break;
        case 'a':
            args.arrayname = optarg;
            break;
        case 'p':
            args.partitioningname = optarg;
            break;
        case 'l':
            args.thread_logname = optarg;
            break;
        case 'v':
            args.max_VV = std::atoi(optarg);
            if (args.max_VV <= 0) {
                std::cerr << "Max VV must be >= 1\n";
                bad++;
            } else if (args.max_VV % 32 != 0) {
                std::cerr << "Max VV is preferred to be a multiple of 32!\n";
            }
            break;
        case 'g':
            args.n_GPUS = std::atoi(optarg);
            if (args.n_GPUS < 0) {
                std::cerr << "GPU count must be non-negative\n";
                bad++;
            }
            break;
        case 'd':
            args.debug = std::atoi(optarg);
            if (args.debug != NO_DEBUG && args.debug != DEBUG_MIN && args.debug != DEBUG_MAX) {
                std::cerr << "Invalid debug level specified. Must be within:\n"
                             "\t0 - No debug\n\t1 - Reduced debug\n\t2 - Verbose debug\n";
                bad++;
            }
            break;
        case 'n':
            args.no_partitioning = true;
            break;
        case 'h':
            std::cout << usage(argv[0], long_opts, help_info, metavars);
            std::exit(EXIT_SUCCESS);
        }
    }
    // one GPU, no partitioning across devices
    if (args.no_partitioning) {
        if (args.n_GPUS > 1) std::cerr << "When partitioning is disabled, only one GPU can be utilized\n";
        args.n_GPUS = 1;
        args.partitioningname = "";
    } else if (args.partitioningname.empty()) {
        std::cerr << "No partitioning name given, treating as --no_partitioning\n";
        args.no_partitioning = true;
        args.n_GPUS = 1;
    }
    int visible = 0;
    tg_device_count(&visible);
    if (args.n_GPUS == 0) {
        if (visible == 0) {
            errors << "No GPUs detected on the system\n";
            bad++;
        }
        args.n_GPUS = visible;
    } else if (visible < args.n_GPUS) {
        std::cerr << "Fewer GPUs detected on the system (" << visible << ") than requested ("
                  << args.n_GPUS << ")\n";
    }
    if (args.n_GPUS > 1) {
        std::cerr << "This build runs on one GPU; using GPU 0\n";
        args.n_GPUS = 1;
    }
    if (args.fileName.empty()) {
        errors << "Must supply an input filename via -i | --input\n";
        bad++;
    }
    if (args.debug >= DEBUG_MIN) {
        std::cout << "Debug level: "
                  << (args.debug == NO_DEBUG ? "None"
                                             : args.debug == DEBUG_MIN ? "Reduced debug"
                                                                       : "Verbose debug")
                  << "\nDataset: " << args.fileName << "\nCPU threads per GPU: "
                  << args.threadNumber << "\nExport: " << (args.export_.empty() ? "[n/a]" : args.export_)
                  << "\nArray name: "
                  << (args.arrayname.empty() ? "[default to first array]" : args.arrayname)
                  << "\nMax VV: "
                  << (args.max_VV == -1 ? "[Estimated from mesh]" : std::to_string(args.max_VV))
                  << "\nGPUs: " << args.n_GPUS << "\n";
    }
    for (int b = 0; b < 14; b++)
        if (!args.arg_flags[b]) args.arg_flags[b] = bits[b] != 0;
    if (no_validate) args.arg_flags[0] = false;
    if (export_vertices) args.export_vertices = true;
    if (args.debug >= DEBUG_MIN)
        for (int b = 0; b < 14; b++)
            std::cout << "Bit flag [" << b << ": " << args.flag_names[b]
                      << "]: " << (args.arg_flags[b] ? "true" : "false") << "\n";
    if (bad) {
        std::cerr << errors.str();
        std::exit(EXIT_FAILURE);
    }
}
