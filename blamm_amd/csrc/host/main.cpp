// `blamm` command-line entry: dict | hist | scan (blstools.cpp:113-166).
#include <iostream>
#include <stdexcept>
#include <string>

#include "host.hpp"

namespace {

void usage()
{
    std::cout << "Usage: blamm command [options]\n\n"
                 " command\n"
                 "  dict\t\t\tmake a dictionary for the input sequences\n"
                 "  hist\t\t\tgenerate PWM score histograms\n"
                 "  scan\t\t\tscan for pwm occurrences (MI355X)\n\n"
                 " [options]\n"
                 "  -h\t--help\t\tdisplay help page\n"
                 "  -v\t--version\tdisplay version\n\n";
}

}  // namespace

int main(int argc, char** argv)
{
    const std::string cmd = argc > 1 ? argv[1] : "";
    if (cmd == "-h" || cmd == "--help") {
        usage();
        return EXIT_SUCCESS;
    }
    if (cmd == "-v" || cmd == "--version") {
        std::cout << "BLAS Accelerated Motif Matching (blamm) -- MI355X scan engine 1.0.0\n";
        return EXIT_SUCCESS;
    }
    int rc;
    try {
        if (cmd == "dict") rc = blamm::dict_main(argc, argv);
        else if (cmd == "hist") rc = blamm::hist_main(argc, argv);
        else if (cmd == "scan") rc = blamm::scan_main(argc, argv);
        else {
            usage();
            return EXIT_FAILURE;
        }
    } catch (const std::exception& e) {
        // runtime errors end the run with a message on stderr (blstools.cpp:93-101)
        std::cerr << e.what() << std::endl;
        return EXIT_FAILURE;
    }
    if (rc == EXIT_SUCCESS) std::cout << "Exiting... bye!" << std::endl;
    return rc;
}
