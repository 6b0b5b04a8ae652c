// main.cpp -- `axs_main <ParameterFile>`: drop-in for the reference's
// ./main (src/main.cpp:3-27) on the standard (uniform-lattice) path.
#include <stdio.h>
#include <stdlib.h>

#include "axs_host.h"

int main(int argc, char *argv[])
{
    if (argc != 2) {
        fprintf(stdout, "Error: usage ./main <ParameterFile>\n");      // main.cpp:5-8
        return EXIT_FAILURE;
    }
    return axs_host_run(argv[1]) == 0 ? EXIT_SUCCESS : EXIT_FAILURE;
}
