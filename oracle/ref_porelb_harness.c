/*
 * Harness around the reference solver, compiled from the reference's own
 * porelb.c where it lies (oracle/Makefile puts that directory on the include
 * path; nothing is copied).  The reference main (porelb.c:402-443) hard-codes
 * tau = 0.8 and Xforce = 1e-3 in init_para (porelb.c:83-94); this harness
 * renames that main and runs the same sequence of reference calls with tau,
 * the body force and the save interval taken from the command line:
 *
 *   alloc_memory, check_V, init_para (then the overrides),
 *   build_send_recv_info, init_dist_macro, and per step
 *   update_send_dist, send_recv, stream_inner_nodes, MPI_Waitall x2,
 *   update_recv_dist, collision, the rank-0 step print, save_data at
 *   it % it_save_data == 0; final save_data.  MPI_Wtime brackets the loop
 *   as in the reference (porelb.c:424-439).
 *
 * Optionally every rank dumps its f0 and f1 (AoS, N*19 doubles) to
 * <prefix>_f0_PPP.bin / <prefix>_f1_PPP.bin after the loop.
 *
 * TEST INFRASTRUCTURE ONLY: produces golden fixtures, pins the oracle, and is
 * the CPU baseline that bench.py times.
 *
 * usage: mpirun -np P ref_porelb.x it_max tau gx save_every print(0|1) [dump_prefix]
 * (reads V.bin from the working directory, like the reference)
 */
#define main plb_reference_porelb_main
#include "porelb.c"
#undef main

int main(int argc, char *argv[])
{
    SIMULATION sim;
    MPI_Init(&argc, &argv);
    if (argc != 6 && argc != 7) {
        fprintf(stderr, "usage: %s it_max tau gx save_every print [dump_prefix]\n", argv[0]);
        exit(-10);
    }
    sim.it_max = atoi(argv[1]);
    double tau = atof(argv[2]);
    double gx = atof(argv[3]);
    int save_every = atoi(argv[4]);
    int print = atoi(argv[5]);

    MPI_Comm_rank(MPI_COMM_WORLD, &sim.pid);
    MPI_Comm_size(MPI_COMM_WORLD, &sim.num_p);

    alloc_memory(&sim);
    check_V(&sim);
    init_para(&sim);
    sim.tau = tau;
    sim.omega = 1.0 / sim.tau;
    sim.nu = 1.0 / 3.0 * (sim.tau - 0.5);
    sim.Xforce = gx;
    sim.it_save_data = save_every > 0 ? save_every : 1000;
    build_send_recv_info(&sim);
    init_dist_macro(&sim);

    double t1, t2;
    t1 = MPI_Wtime();
    for (sim.it = 0; sim.it < sim.it_max; sim.it++) {
        update_send_dist(&sim);
        send_recv(&sim);
        stream_inner_nodes(&sim);
        MPI_Waitall(sim.N_p, sim.send_request, MPI_STATUS_IGNORE);
        MPI_Waitall(sim.N_p, sim.recv_request, MPI_STATUS_IGNORE);
        update_recv_dist(&sim);
        collision(&sim);
        if (print && !sim.pid) printf(">>> Step : %20d \n", sim.it);
        if (save_every >= 0 && sim.it % sim.it_save_data == 0) save_data(&sim);
    }
    t2 = MPI_Wtime();
    long long nloc = sim.N, ntot = 0;
    MPI_Reduce(&nloc, &ntot, 1, MPI_LONG_LONG, MPI_SUM, 0, MPI_COMM_WORLD);
    if (!sim.pid) {
        printf("Elapsed time is %f\n", t2 - t1);
        printf("PLB_REF_RESULT nodes=%lld steps=%d ranks=%d seconds=%.9f mlups=%.6f\n", ntot,
               sim.it_max, sim.num_p, t2 - t1,
               (t2 > t1) ? (double)ntot * sim.it_max / (t2 - t1) / 1e6 : 0.0);
    }
    if (save_every >= 0) save_data(&sim);
    if (argc == 7) {
        char fn[512];
        FILE *fp;
        snprintf(fn, sizeof fn, "%s_f0_%03d.bin", argv[6], sim.pid);
        fp = fopen(fn, "wb");
        fwrite(sim.f0, sizeof(double), (size_t)sim.N * Q, fp);
        fclose(fp);
        snprintf(fn, sizeof fn, "%s_f1_%03d.bin", argv[6], sim.pid);
        fp = fopen(fn, "wb");
        fwrite(sim.f1, sizeof(double), (size_t)sim.N * Q, fp);
        fclose(fp);
    }
    MPI_Finalize();
    return 0;
}
