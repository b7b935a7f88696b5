"""Worker of the full-batch GPU test (not a test module): one system of the
uneven 64-system workload and its FP64 sparse direct solution, built in a
spawned CPU process so that the 64 direct solves run side by side."""


def build_and_solve(spec):
    from hsc_reimpl_amd import systems
    from oracle import sparse as osp
    s = systems.batch_system(spec)
    return s, osp.direct_solve(s)
