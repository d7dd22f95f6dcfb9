"""ctypes mirror of include/mbl.h (structures and constants only).

The structures follow the reference data model (lbm/data_type_lbm.h:7-51,
multi/data_type_multi.h:7-31); device pointers are carried as c_void_p, which
has the size and alignment of any C pointer, so the layouts match the header
for both library builds.
"""
import ctypes as C

Q = 9
MAX_LEVEL = 8

FLUID = 0
SOLID_SBB = 1
SOLID_WALL = 2
BOUNDARY_INLET = 5
BOUNDARY_OUTLET = 7
C_BOUNDARY = 9

STORE_NATIVE, STORE_F32S, STORE_F16S, STORE_BF16S, STORE_F16C, STORE_F16D = 0, 1, 2, 3, 4, 5
BC_REFERENCE, BC_CHANNEL, BC_PERIODIC = 0, 1, 2
INIT_UNIFORM, INIT_REST, INIT_RANDOM, INIT_PROFILE = 0, 1, 2, 3

E_NAMES = {0: "ok", -1: "invalid", -2: "hip", -3: "nomem", -4: "not finite",
           -5: "diverged", -6: "no handle", -7: "unsupported"}


class P4(C.Structure):
    _fields_ = [("xstart", C.c_int), ("xend", C.c_int), ("ystart", C.c_int), ("yend", C.c_int)]


class P3i(C.Structure):
    _fields_ = [("nc", C.c_int), ("nf", C.c_int), ("direction", C.c_int)]


class P4d(C.Structure):
    _fields_ = [("xstart", C.c_double), ("xend", C.c_double), ("ystart", C.c_double), ("yend", C.c_double)]


class P4d4i(C.Structure):
    _fields_ = [("reduced", P4d), ("move", P4)]


class Int3(C.Structure):
    _fields_ = [("i", C.c_int), ("j", C.c_int), ("k", C.c_int)]


class AREA(C.Structure):
    _fields_ = [("m", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("base_x", C.c_double),
                ("base_y", C.c_double), ("tau", C.c_double), ("rv_tau", C.c_double)]


class Inf_Ff_store(C.Structure):
    _fields_ = [("areac_n", C.c_int), ("Nc", C.c_int), ("line_C", P4),
                ("areaf_n", C.c_int), ("Nf", C.c_int), ("line_F", P4)]


class D2Q9(C.Structure):
    _fields_ = [("solid", C.c_void_p), ("x", C.c_void_p), ("y", C.c_void_p), ("rho", C.c_void_p),
                ("ux", C.c_void_p), ("uy", C.c_void_p), ("forcex", C.c_void_p), ("forcey", C.c_void_p),
                ("f", C.c_void_p * Q), ("F", C.c_void_p * Q)]


class F_k3(C.Structure):
    _fields_ = [("F", C.c_void_p * Q)]


class Ff_Store(C.Structure):
    _fields_ = [("N", C.c_int), ("marker", C.c_int * 3), ("f_t1", F_k3), ("f_t", F_k3), ("f_t_1", F_k3)]


class FM_spline3(C.Structure):
    _fields_ = [("N", C.c_int), ("l", C.c_void_p), ("m", C.c_void_p), ("u", C.c_void_p),
                ("Varibs_M", C.c_void_p), ("rho", C.c_void_p), ("ux", C.c_void_p), ("uy", C.c_void_p),
                ("F", C.c_void_p * Q)]


class mbl_dim3(C.Structure):
    _fields_ = [("x", C.c_uint), ("y", C.c_uint), ("z", C.c_uint)]


class mbl_config(C.Structure):
    _fields_ = [("LEVEL", C.c_int), ("mi", C.c_int * MAX_LEVEL), ("number_line2", C.c_int),
                ("rho0", C.c_double), ("U0", C.c_double * 2), ("LX", C.c_int), ("LY", C.c_int),
                ("R", C.c_double), ("CX", C.c_double), ("CY", C.c_double),
                ("bc_mode", C.c_int), ("init_mode", C.c_int), ("storage", C.c_int),
                ("seed", C.c_ulonglong), ("sigma", C.c_double),
                ("use_graph", C.c_int), ("sync_diag", C.c_int), ("flush_interval", C.c_int),
                ("write_files", C.c_int), ("out_dir", C.c_char * 256), ("number_area", C.c_int)]


class mbl_geometry(C.Structure):
    _fields_ = [("LEVEL", C.c_int), ("mi", C.c_int * MAX_LEVEL),
                ("number_area", C.c_int), ("number_line", C.c_int), ("number_line2", C.c_int),
                ("ASP", C.c_int * (MAX_LEVEL + 1)), ("LTCFFCP", C.c_int * (MAX_LEVEL + 1)),
                ("LS3CFFCP", C.c_int * (MAX_LEVEL + 1)), ("LCCP", C.c_int * (MAX_LEVEL + 1)),
                ("AreaZone", C.POINTER(P4d4i)), ("AreaDefine", C.POINTER(P4)),
                ("NeborCF", C.POINTER(P3i)), ("NeborCC", C.POINTER(P3i)),
                ("LX", C.c_int), ("LY", C.c_int), ("Re", C.c_double), ("U0_2", C.c_double), ("R", C.c_double)]


class mbl_case_spec(C.Structure):
    _fields_ = [("LX", C.c_int), ("LY", C.c_int), ("region", P4), ("R", C.c_double), ("CX", C.c_double),
                ("CY", C.c_double), ("Re", C.c_double), ("U", C.c_double)]


class mbl_channel_spec(C.Structure):
    _fields_ = [("LX", C.c_int), ("LY", C.c_int), ("levels", C.c_int), ("region", P4 * MAX_LEVEL),
                ("Re", C.c_double), ("U", C.c_double), ("R", C.c_double)]


# Exported symbols of include/mbl.h (both builds export the same names).
EXPORTS = [
    "mbl_config_default", "mbl_handle_create", "mbl_handle_destroy", "mbl_handle_make_current",
    "mbl_handle_current", "mbl_get_last_error", "mbl_error_string", "mbl_synchronize",
    "mbl_flush_diagnostics", "mbl_set_stream", "mbl_malloc", "mbl_free", "mbl_memcpy_h2d",
    "mbl_memcpy_d2h", "mbl_memcpy_d2d", "mbl_memset_d", "mbl_plane_bytes", "mbl_block_alloc",
    "mbl_block_free", "mbl_read_distribution", "mbl_write_distribution", "mbl_geo_init",
    "mbl_geometry_preset", "mbl_geometry_channel", "init_General", "init_Fstore", "init_tri_A",
    "evolution", "sum_N_drag", "fdrag_store", "getDrag_block", "error_L2", "mbl_last_diagnostics",
    "mbl_diagnostics_series", "mbl_diagnostics_clear", "mbl_divergence_step", "mbl_error_L2_noalloc", "mbl_cycle_launches", "mbl_cycle_positions", "mbl_set_cycle_fork", "mbl_step_block", "mbl_timer_begin",
    "mbl_timer_end", "mbl_set_tuning", "mbl_poiseuille_l2", "mbl_set_persistent",
    "mbl_batch_create", "mbl_batch_destroy", "mbl_batch_step", "mbl_batch_set_persistent", "mbl_batch_node_updates", "mbl_batch_forces", "mbl_batch_residual",
    "mbl_batch_autotune", "mbl_batch_block", "mbl_batch_raster", "mbl_batch_init_from_raster",
    "mbl_net_create", "mbl_net_destroy", "mbl_net_num_params", "mbl_net_params", "mbl_net_forward",
    "mbl_net_loss_grad", "mbl_net_train", "output", "mbl_output", "mbl_checkpoint_save", "mbl_checkpoint_load",
    "mbl_bandwidth_probe", "mbl_bandwidth_probe_planes",
]
