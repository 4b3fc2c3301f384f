// Grid size of the giant-island kernels (k_t_giant, k_t_color_giant). Their workgroups meet at
// grid-wide barriers, so every workgroup of the grid must be able to be resident at the same time:
// the grid is the preferred size, lowered to what the occupancy of the two kernels allows on the
// device's compute units. Plain C++ (no HIP types) so that the decision is testable on a CPU.
#ifndef ME_GIANT_GRID_HPP
#define ME_GIANT_GRID_HPP

namespace me {

// preferred: the grid the kernels are tuned for; per_cu_sweep / per_cu_color: workgroups of each
// kernel that one compute unit holds at once (hipOccupancyMaxActiveBlocksPerMultiprocessor);
// n_cu: compute units. Returns the grid, or 0 when not even one workgroup fits (the engine then
// refuses an island beyond one workgroup instead of launching).
inline int giant_grid_size(int preferred, int per_cu_sweep, int per_cu_color, int n_cu) {
    if (preferred <= 0 || per_cu_sweep <= 0 || per_cu_color <= 0 || n_cu <= 0) return 0;
    long long cap_s = (long long)per_cu_sweep * n_cu, cap_c = (long long)per_cu_color * n_cu;
    long long g = preferred;
    if (g > cap_s) g = cap_s;
    if (g > cap_c) g = cap_c;
    return (int)g;
}

}  // namespace me

#endif
