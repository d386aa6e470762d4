// main.cpp -- the solver executable: muphfasa <inputs> [key=value ...]
#include <cstdio>

#include "../../include/mfa_host.h"

int main(int argc, char* argv[])
{
    const int rc = mfa_run_main(argc, (const char* const*)argv);
    if (rc != 0) {
        std::fprintf(stderr, "muphfasa: %s\n", mfa_sim_error());
        return 1;
    }
    return 0;
}
