/* scalfmm_capi.h -- the reference's public C API (Addons/CKernelApi/Src/
 * CScalfmmApi.h) for its built-in Chebyshev kernel, served by the MI355X
 * engine of scalfmm_hip.h.
 *
 * The names, argument lists and enum values are the reference's, so a C or
 * Fortran program written against CScalfmmApi.h for `chebyshev` links against
 * this library unchanged.  The kernel is FChebSymKernel of order 7 with the
 * 1/r matrix kernel in FP64, as the reference's scalfmm_init builds it
 * (Addons/CKernelApi/Src/FScalfmmApiInit.cpp:37-46, 79-89).  Supported:
 * kernel type `chebyshev`; algorithms `sequential`, `multi_thread` (one
 * engine, the same results) and `source_target` (sources and targets as
 * separate sets).  Not supported (the functions are absent): user-defined
 * kernels and their callbacks, the periodic, adaptive and MPI algorithms,
 * tree_abstract_insert.  Where the reference aborts (FAssertLF, exit), these
 * functions print a message on stderr and return.
 *
 * Particles are addressed by their index within the insert call that added
 * them, as in the reference (FInterEngine.hpp:96-140).  Results accumulate
 * into the particles' potentials and forces at every execution, which start
 * at zero on insertion (set_potentials / set_forces reset them). */
#ifndef SCALFMM_CAPI_H
#define SCALFMM_CAPI_H

#ifdef __cplusplus
extern "C" {
#endif

/* CScalfmmApi.h:49-57 */
typedef enum kernel_type {
    user_defined_kernel = 0,
    chebyshev = 1,
    lagrange = 2
} scalfmm_kernel_type;

/* CScalfmmApi.h:77-84 */
typedef enum scalfmm_algorithm_config {
    sequential = 0,
    multi_thread = 1,
    periodic = 2,
    source_target = 3,
    adaptiv = 4,
    mpi = 5
} scalfmm_algorithm;

/* CScalfmmApi.h:269-273 */
typedef enum particule_type {
    SOURCE = 0,
    TARGET = 1,
    BOTH = 2
} PartType;

typedef void* scalfmm_handle;

/* The cell and leaf descriptors of user-defined kernels (CScalfmmApi.h:231-248):
 * five callbacks each; ignored by the Chebyshev kernel (pass zeroed structs). */
typedef struct User_Scalfmm_Cell_Descriptor {
    void* user_free_cell;
    void* user_init_cell;
    void* user_get_size;
    void* user_copy_cell;
    void* user_restore_cell;
} Scalfmm_Cell_Descriptor;
typedef struct User_Scalfmm_Leaf_Descriptor {
    void* user_init_leaf;
    void* user_free_leaf;
    void* user_get_size;
    void* user_copy_leaf;
    void* user_restore_leaf;
} Scalfmm_Leaf_Descriptor;
typedef void (*Callback_free_cell)(void*);

/* CScalfmmApi.h:103: NULL for kernel types other than chebyshev or for the
 * periodic, adaptive and MPI algorithms. */
scalfmm_handle scalfmm_init(scalfmm_kernel_type KernelType, scalfmm_algorithm algo);
/* CScalfmmApi.h:259-262: creates the engine (order 7) for this box. */
void scalfmm_build_tree(scalfmm_handle handle, int TreeHeight, double BoxWidth, double* BoxCenter,
                        Scalfmm_Cell_Descriptor user_cell_descriptor, Scalfmm_Leaf_Descriptor user_leaf_descriptor);
/* CScalfmmApi.h:317-324 */
void scalfmm_tree_insert_particles(scalfmm_handle Handle, int NbPositions, double* arrayX, double* arrayY,
                                   double* arrayZ, PartType type);
void scalfmm_tree_insert_particles_xyz(scalfmm_handle Handle, int NbPositions, double* XYZ, PartType type);
/* CScalfmmApi.h:339-365 */
void scalfmm_set_physical_values(scalfmm_handle Handle, int nbPhysicalValues, double* physicalValues, PartType type);
void scalfmm_get_physical_values(scalfmm_handle Handle, int nbPhysicalValues, double* physicalValues, PartType type);
void scalfmm_set_physical_values_npart(scalfmm_handle Handle, int nbPhysicalValues, int* idxOfParticles,
                                       double* physicalValues, PartType type);
void scalfmm_get_physical_values_npart(scalfmm_handle Handle, int nbPhysicalValues, int* idxOfParticles,
                                       double* physicalValues, PartType type);
/* CScalfmmApi.h:383-405 (forces of sources have no meaning: the reference prints so) */
void scalfmm_get_forces_xyz(scalfmm_handle Handle, int nbParts, double* forcesToFill, PartType type);
void scalfmm_get_forces_xyz_npart(scalfmm_handle Handle, int nbParts, int* idxOfParticles, double* forcesToFill,
                                  PartType type);
void scalfmm_get_forces(scalfmm_handle Handle, int nbParts, double* fX, double* fY, double* fZ, PartType type);
void scalfmm_get_forces_npart(scalfmm_handle Handle, int nbParts, int* idxOfParticles, double* fX, double* fY,
                              double* fZ, PartType type);
void scalfmm_set_forces_xyz(scalfmm_handle Handle, int nbParts, double* forcesToFill, PartType type);
void scalfmm_set_forces_xyz_npart(scalfmm_handle Handle, int nbParts, int* idxOfParticles, double* forcesToFill,
                                  PartType type);
void scalfmm_set_forces(scalfmm_handle Handle, int nbParts, double* fX, double* fY, double* fZ, PartType type);
void scalfmm_set_forces_npart(scalfmm_handle Handle, int nbParts, int* idxOfParticles, double* fX, double* fY,
                              double* fZ, PartType type);
/* CScalfmmApi.h:420-423 */
void scalfmm_get_potentials(scalfmm_handle Handle, int nbParts, double* potentialsToFill, PartType type);
void scalfmm_set_potentials(scalfmm_handle Handle, int nbParts, double* potentialsToRead, PartType type);
void scalfmm_get_potentials_npart(scalfmm_handle Handle, int nbParts, int* idxOfParticles, double* potentialsToFill,
                                  PartType type);
void scalfmm_set_potentials_npart(scalfmm_handle Handle, int nbParts, int* idxOfParticles, double* potentialsToRead,
                                  PartType type);
/* CScalfmmApi.h:436-472: positions; moved particles take effect at the next
 * execution (the engine rebuilds its tree every time; scalfmm_update_tree
 * is accepted for compatibility). */
void scalfmm_add_to_positions_xyz(scalfmm_handle Handle, int NbPositions, double* updatedXYZ, PartType type);
void scalfmm_add_to_positions(scalfmm_handle Handle, int NbPositions, double* X, double* Y, double* Z, PartType type);
void scalfmm_add_to_positions_xyz_npart(scalfmm_handle Handle, int NbPositions, int* idxOfParticles,
                                        double* updatedXYZ, PartType type);
void scalfmm_add_to_positions_npart(scalfmm_handle Handle, int NbPositions, int* idxOfParticles, double* X,
                                    double* Y, double* Z, PartType type);
void scalfmm_set_positions_xyz(scalfmm_handle Handle, int NbPositions, double* updatedXYZ, PartType type);
void scalfmm_set_positions(scalfmm_handle Handle, int NbPositions, double* X, double* Y, double* Z, PartType type);
void scalfmm_set_positions_xyz_npart(scalfmm_handle Handle, int NbPositions, int* idxOfParticles, double* updatedXYZ,
                                     PartType type);
void scalfmm_set_positions_npart(scalfmm_handle Handle, int NbPositions, int* idxOfParticles, double* X, double* Y,
                                 double* Z, PartType type);
void scalfmm_update_tree(scalfmm_handle handle);
void scalfmm_get_positions_xyz(scalfmm_handle Handle, int NbPositions, double* positionsToFill, PartType type);
void scalfmm_get_positions(scalfmm_handle Handle, int NbPositions, double* X, double* Y, double* Z, PartType type);
void scalfmm_get_positions_xyz_npart(scalfmm_handle Handle, int NbPositions, int* idxOfParticles,
                                     double* positionsToFill, PartType type);
void scalfmm_get_positions_npart(scalfmm_handle Handle, int NbPositions, int* idxOfParticles, double* X, double* Y,
                                 double* Z, PartType type);
/* CScalfmmApi.h:491: sequential, multi_thread or source_target (the latter
 * only before particles are inserted). */
void scalfmm_algorithm_config(scalfmm_handle Handle, scalfmm_algorithm config);
/* CScalfmmApi.h:738-745: all passes, or P2M/M2M/M2L/L2L/L2P only. */
void scalfmm_execute_fmm(scalfmm_handle Handle);
void scalfmm_execute_fmm_far_field(scalfmm_handle Handle);
/* CScalfmmApi.h:772 */
void scalfmm_dealloc_handle(scalfmm_handle handle, Callback_free_cell cellDestroyer);
/* CScalfmmApi.h:790-802: 9 timers, of which the first 7 are filled (seconds of
 * the last execution: P2M, M2M, M2L, L2L, L2P, P2P, near field), as
 * FScalFMMEngine::get_timers (FScalFMMEngine.hpp:782-794) fills them. */
int scalfmm_get_nb_timers(scalfmm_handle handle);
void scalfmm_get_timers(scalfmm_handle handle, double* Timers);
/* CScalfmmApi.h:816: not available with the Chebyshev kernel (the reference
 * aborts; this prints so). */
void scalfmm_set_upper_limit(scalfmm_handle handle, int upperLimit);

#ifdef __cplusplus
}
#endif
#endif
