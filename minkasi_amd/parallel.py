"""The reference's MPI bootstrap names (minkasi/parallel.py:5-79) for one GPU.

Every TOD of a job lives on the one device, so the process is rank 0 of 1:
there is no MPI communicator, barrier() only waits for the device, and the
OpenMP thread controls have no counterpart (set_nthread does nothing,
get_nthread reports the host threads this process may use)."""
import os

MPI = None
comm = None
have_mpi = False
myrank = 0
nproc = 1

__all__ = ["MPI", "have_mpi", "comm", "myrank", "nproc", "report_mpi", "barrier", "set_nthread", "get_nthread"]


def report_mpi():
    """parallel.py:41-48."""
    print("mpi not found")


def barrier():
    """parallel.py:50-58: waits for the device's queued work."""
    import torch
    if torch.cuda.is_available() and torch.cuda.is_initialized():
        torch.cuda.synchronize()


def set_nthread(nthread):
    """parallel.py:61-70 (OpenMP threads of the C library): nothing to set."""
    return None


def get_nthread():
    """parallel.py:73-79."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1
