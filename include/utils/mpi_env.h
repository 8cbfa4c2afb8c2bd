// MPI set-up of the driver (mpi_env.h of the reference): MPI with
// MPI_THREAD_FUNNELED for the lifetime of the object.  The one-GPU build runs
// one rank per box; the drop-in classes reject communicators of other sizes.
#pragma once
#include <mpi.h>

class MPIEnv {
public:
    MPIEnv(int& argc, char**& argv) {
        int provided = 0;
        MPI_Init_thread(&argc, &argv, MPI_THREAD_FUNNELED, &provided);
    }
    ~MPIEnv() { MPI_Finalize(); }
    MPIEnv(const MPIEnv&) = delete;
    MPIEnv& operator=(const MPIEnv&) = delete;
};
