/*
 * argparse.h -- command line of the reference driver (src/argparse.cu:52-326):
 * the same short and long options, help texts, defaults and flag bits
 * (-h -i -t -e -a -p -l -v -g -d -n, --validate, --build_TE ... --build_VV).
 * Flags can only be switched on; without --validate support compiled in,
 * bit 0 keeps its default, as in the reference.  Invalid arguments print
 * the collected errors and exit(EXIT_FAILURE); -h prints usage and exits.
 */
#ifndef TETRA_ARGPARSE_H
#define TETRA_ARGPARSE_H

#include <getopt.h>

#include <string>

#include "datatypes.h"

std::string usage(const char *argv0, const struct option *options, const option_map &help_info,
                  const option_map &metavars);
void parse(int argc, char *argv[], runtime_arguments &args);

#endif
