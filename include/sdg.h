/*
 * sdg.h -- C-ABI of libsdg_hip.so, the MI355X (gfx950) discontinuous Galerkin residual and
 * explicit SSP-RK time-stepping engine.
 *
 * The reference has no FFI: its operator boundary is the C++ template Solver<SimulationControl>
 * (src/Solver/SolveControl.cpp:377-435), driven by System<SimulationControl>::solve
 * (src/Utils/SystemControl.cpp:159-195). Every entry point below replaces one member of that
 * template (cited per function), so a C++ System<SC> adaptor, a ctypes test harness or a
 * torch.autograd wrapper can drive the same hot path.
 *
 * Conventions
 *   - every function returns SDG_OK (0) or a negative SDG_E_* code and never throws;
 *     sdg_last_error() returns the message of the last failure in a context (or globally).
 *   - the library owns all device memory; callers own host buffers, which are borrowed for the
 *     duration of the call only.
 *   - one host thread per context; distinct contexts are independent (no process-global
 *     physical state, unlike the reference's inline static members, PhysicalModel.cpp:28-29).
 *   - coefficient arrays in the "reference layout" are, per element, the NV x N_b matrix of
 *     modal coefficients stored column-major (variable index fastest), elements in mesh order,
 *     exactly as Eigen stores variable_basis_function_coefficient_ (SolveControl.cpp:47-57).
 *   - conserved variables are (rho, rho*u, rho*v[, rho*w], rho*E) (VariableConvertor.cpp:28-48);
 *     primitive variables are (rho, u, v[, w], T) (VariableConvertor.cpp:118-139).
 */
#ifndef SDG_H_
#define SDG_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SDG_ABI_VERSION 5

/* ---- status codes ---------------------------------------------------------------------- */
#define SDG_OK 0
#define SDG_E_INVALID (-1)     /* bad argument or descriptor                              */
#define SDG_E_UNSUPPORTED (-2) /* valid request this build does not implement             */
#define SDG_E_DEVICE (-3)      /* HIP runtime error (no device, launch failure, ...)      */
#define SDG_E_NOMEM (-4)       /* host or device allocation failed                        */
#define SDG_E_STATE (-5)       /* call out of order (e.g. step before a state was set)    */

/* ---- enumerations mirroring src/Utils/Enum.cpp:22-233 ----------------------------------- */
enum sdg_element {
  SDG_ELEM_LINE = 1,
  SDG_ELEM_TRIANGLE = 2,
  SDG_ELEM_QUADRANGLE = 3,
  SDG_ELEM_TETRAHEDRON = 4,
  SDG_ELEM_PYRAMID = 5,
  SDG_ELEM_HEXAHEDRON = 6,
  SDG_ELEM_MIXED_TRIANGLE_QUADRANGLE = 100 /* mesh info of a split grid holding both types */
};
enum sdg_equation { SDG_EQ_COMPRESSIBLE_EULER = 0, SDG_EQ_COMPRESSIBLE_NS = 1 };
enum sdg_convective_flux {
  SDG_FLUX_CENTRAL = 0,
  SDG_FLUX_LAX_FRIEDRICHS = 1,
  SDG_FLUX_HLLC = 2,
  SDG_FLUX_ROE = 3,
  SDG_FLUX_EXACT = 4
};
enum sdg_viscous_flux { SDG_VISC_NONE = 0, SDG_VISC_BR1 = 1, SDG_VISC_BR2 = 2 };
enum sdg_time_integration { SDG_RK_FORWARD_EULER = 0, SDG_RK_HEUN2 = 1, SDG_RK_SSPRK3 = 2 };
enum sdg_eos { SDG_EOS_IDEAL_GAS = 0, SDG_EOS_WEAK_COMPRESSIBLE = 1 };
enum sdg_transport { SDG_TRANSPORT_NONE = 0, SDG_TRANSPORT_CONSTANT = 1, SDG_TRANSPORT_SUTHERLAND = 2 };
enum sdg_precision { SDG_PREC_FP64 = 0, SDG_PREC_MIXED = 1 };
/* SourceTermEnum (Enum.cpp:71-74), same order. Boussinesq: rho beta (T - T_ref) g, g = 1, on the
 * last momentum component (y in 2-D, z in 3-D), added to the residual as int S phi
 * (SourceTerm.cpp:29-58, SpatialDiscrete.cpp:254-262,1026-1029). */
enum sdg_source_term { SDG_SOURCE_NONE = 0, SDG_SOURCE_BOUSSINESQ = 1 };
/* ShockCapturingEnum (Enum.cpp:76-79), same order. Artificial viscosity (Persson-Peraire sensor on
 * the density, SpatialDiscrete.cpp:37-192, ViscousFlux.cpp:105-190): once per step, per element
 * s = log10(|rho_high|^2 / |rho|^2) over the modes above order P-1, eps_e = f r / P times 0, 1 or
 * (1 + sin(pi (s - s0) / (2 k))) / 2 below, above or inside [s0 - k, s0 + k], s0 = log10(1 / P^4)
 * (0 for P = 1), k the empirical tolerance, f the factor, r the inner radius; the node value is
 * the maximum over the elements that share the node, and the P1 interpolant of the node values
 * times the volume gradient of the conserved variables is subtracted from the fluxes (central
 * average on interior faces, the element's own value on boundary faces). Euler and Navier-Stokes
 * (single problems; with Navier-Stokes the viscous flux keeps its own lifted gradient), in FP64 and,
 * for Euler on the matrix-core sizes and Navier-Stokes on hexahedra P2-P4, in the mixed mode. */
enum sdg_shock_capturing { SDG_SHOCK_NONE = 0, SDG_SHOCK_ARTIFICIAL_VISCOSITY = 1 };
/* InitialConditionEnum (Enum.cpp), same order: the file-based starts of sdg_raw_binary_read */
enum sdg_initial_condition { SDG_IC_FUNCTION = 0, SDG_IC_SPECIFIC_FILE = 1, SDG_IC_LAST_STEP = 2 };
/* BoundaryConditionEnum (Enum.cpp:92-100), same order */
enum sdg_boundary {
  SDG_BC_RIEMANN_FARFIELD = 0,
  SDG_BC_VELOCITY_INFLOW = 1,
  SDG_BC_PRESSURE_OUTFLOW = 2,
  SDG_BC_ISOTHERMAL_NONSLIP_WALL = 3,
  SDG_BC_ADIABATIC_SLIP_WALL = 4,
  SDG_BC_ADIABATIC_NONSLIP_WALL = 5,
  SDG_BC_PERIODIC = 6
};

/* Kernel family of a single-problem context. AUTO: sum-factorised Navier-Stokes kernels on
 * hexahedra P2-P4, dense matrix-core kernels on the other matrix-core sizes (quadrangles P3-P5,
 * hexahedra P2-P4 Euler), vector kernels elsewhere. DENSE keeps the dense matrix-core kernels on
 * hexahedra; VECTOR runs the vector kernels for every size (FP64 only). All evaluate the same
 * operator; the choice exists to compare them. */
enum sdg_kernels { SDG_KERNELS_AUTO = 0, SDG_KERNELS_DENSE = 1, SDG_KERNELS_VECTOR = 2 };
/* GRAPH: each sdg_step call replays whole steps from captured hipGraphs; EAGER: every kernel is
 * launched on the stream as it comes (same kernels, same results). */
enum sdg_launch { SDG_LAUNCH_GRAPH = 0, SDG_LAUNCH_EAGER = 1 };

/* Runtime copy of the reference's compile-time SimulationControl
 * (SimulationControl.cpp:1197-1279) plus the System setters that carry physical constants
 * (SystemControl.cpp:79-109). gamma = 1.4 and Pr = 0.71 are constants, as in
 * PhysicalModel.cpp:45,91,107. */
typedef struct sdg_problem_desc {
  int abi_version;        /* SDG_ABI_VERSION                                              */
  int equation;           /* enum sdg_equation                                            */
  int convective_flux;    /* enum sdg_convective_flux                                     */
  int viscous_flux;       /* enum sdg_viscous_flux (NONE for Euler)                       */
  int time_integration;   /* enum sdg_time_integration                                    */
  int equation_of_state;  /* enum sdg_eos                                                 */
  int transport;          /* enum sdg_transport (NONE for Euler)                          */
  int precision;          /* enum sdg_precision                                           */
  double cp;              /* specific heat at constant pressure                           */
  double cv;              /* specific heat at constant volume                             */
  double dynamic_viscosity;
  double reference_sound_speed; /* weakly-compressible EOS only                           */
  double reference_density;     /* weakly-compressible EOS only                           */
  int source_term;        /* enum sdg_source_term (FP64, single problems; setSourceTerm,  */
                          /* SystemControl.cpp:65-68)                                     */
  int pad_;
  double thermal_expansion_coefficient; /* Boussinesq beta                                */
  double reference_temperature;         /* Boussinesq T_ref                               */
  int shock_capturing;    /* enum sdg_shock_capturing (setArtificialViscosity,            */
                          /* SystemControl.cpp:105-108)                                   */
  int pad2_;
  double empirical_tolerance;           /* k above                                        */
  double artificial_viscosity_factor;   /* f above (the reference's default is 1)         */
  int kernels;            /* enum sdg_kernels: which kernel family runs the stages (no    */
                          /* counterpart in the reference; AUTO unless comparing families)*/
  int launch;             /* enum sdg_launch: whole steps replayed from hipGraphs or      */
                          /* launched kernel by kernel (no counterpart in the reference)  */
} sdg_problem_desc;

/* Structured mesh on an axis-aligned box: the project-side replacement for the gmsh
 * transfinite (+ setPeriodic) generators of the examples (e.g. periodic_2d_ceuler.cpp:63-94).
 * element: SDG_ELEM_LINE (1-D), SDG_ELEM_QUADRANGLE (2-D) or SDG_ELEM_HEXAHEDRON (3-D). A direction is periodic
 * (needs at least 2 cells) unless bounded[d] != 0; the sides of a bounded direction are boundary
 * faces of type boundary[2 d] (low side) and boundary[2 d + 1] (high side), enum sdg_boundary
 * without SDG_BC_PERIODIC. A zero-initialised descriptor tail means fully periodic. */
typedef struct sdg_structured_desc {
  int element;            /* enum sdg_element                                             */
  int polynomial_order;   /* P = 1..5                                                     */
  int n[3];               /* cells per direction (n[2] ignored in 2-D)                    */
  double lo[3];           /* box lower corner                                             */
  double hi[3];           /* box upper corner                                             */
  int bounded[3];         /* 0: periodic direction                                        */
  int boundary[6];        /* enum sdg_boundary per bounded side                           */
} sdg_structured_desc;

/* Sizes of a mesh / discretisation (ElementTrait, SimulationControl.cpp:1083-1108). */
typedef struct sdg_mesh_info {
  int dimension;          /* D                                                            */
  int element;            /* enum sdg_element                                             */
  int polynomial_order;   /* P                                                            */
  int64_t elements;       /* E                                                            */
  int64_t faces;          /* interior faces F (all faces are interior on a periodic mesh) */
  int basis;              /* N_b                                                          */
  int quadrature;         /* N_q   (volume points, Gauss order 2P)                        */
  int faces_per_element;  /* N_f                                                          */
  int face_quadrature;    /* N_fq  (points per face, Gauss order 2P+1)                    */
  int affine;             /* 1 when every element has a constant Jacobian                 */
  int64_t boundary_faces; /* boundary faces F_b                                           */
  int64_t nodes;          /* mesh nodes (element vertices; periodic images are distinct)  */
  int vertices;           /* vertices per element (kBasicNodeNumber)                      */
  int pad_;
} sdg_mesh_info;

/* Operand tables a mesh exports (all FP64 unless noted). Shapes:
 *   BASIS_VALUE        N_q  x N_b      modal_value_                  BasisFunction.cpp:205-211
 *   BASIS_GRADIENT     (N_q*D) x N_b   modal_gradient_value_ (row q*D+d)       :218-226
 *   BASIS_FACE_VALUE   (N_f*N_fq) x N_b modal_adjacency_value_                 :181-186
 *   BASIS_LSQ_INVERSE  N_b x N_b       modal_least_squares_inverse_            :217
 *   QUAD_WEIGHT        N_q             ElementQuadrature::weight_  Quadrature.cpp:37-54
 *   FACE_QUAD_WEIGHT   N_fq
 *   QUAD_COORD         E x N_q x D     physical volume quadrature points (Geometry.cpp:57)
 *   JACOBIAN_DET_W     E x N_q         |J| w                                     Geometry.cpp:61
 *   JACOBIAN_TIW       E x N_q x D*D   (J^T)^-1 |J| w, column-major D x D        Geometry.cpp:63
 *   MASS_INVERSE       E x N_b x N_b   local_mass_matrix_inverse_                Geometry.cpp:88-100
 *   MIN_EDGE           E               minimum_edge_                             Geometry.cpp:29-42
 *   FACE_NORMAL        F x N_fq x D    outward from the left parent              Geometry.cpp:102-168
 *   FACE_JACOBIAN_W    F x N_fq        |J_f| w                                   Geometry.cpp:69-86
 *   FACE_COORD         F x N_fq x D    physical face quadrature points (left parent's view)
 *   FACE_PARENT (int32)     F x 2      left, right element
 *   FACE_LOCAL  (int32)     F x 2      local face index in left, right parent
 *   FACE_SEQUENCE (int32)   F x N_fq   right-parent point index of left point j
 *                                      (getAdjacencyElementQuadratureSequence,
 *                                       SimulationControl.cpp:381-523)
 *   FACE_ROTATION (int32)   F          rotation index as defined by Adjacency.cpp:169-171
 *   BOUNDARY_FACE_COORD     F_b x N_fq x D   physical boundary face points
 *   BOUNDARY_FACE_NORMAL    F_b x N_fq x D   outward from the element
 *   BOUNDARY_FACE_JACOBIAN_W F_b x N_fq      |J_f| w
 *   BOUNDARY_FACE_PARENT (int32) F_b x 2     element, its local face index
 *   BOUNDARY_FACE_TYPE (int32)   F_b         enum sdg_boundary
 *   INNER_RADIUS       E               inner_radius_ (gmsh "innerRadius", Geometry.cpp:29-42):
 *                                      here D |element| / |boundary|
 *   NODE_TAG (int32)   E x nvert       0-based node of each vertex (node_tag_ - 1)
 *   NODAL_VALUE        N_q x nvert     P1 nodal functions at the volume points  BasisFunction.cpp:198-207
 *   NODAL_FACE_VALUE   (N_f*N_fq) x nvert  ... at the face points in parent coordinates  :161-180
 *   BOUNDARY_FACE_PHYSICAL (int32) F_b  gmsh physical tag of a boundary face of a mesh file (-1 otherwise)
 */
enum sdg_mesh_field {
  SDG_FIELD_BASIS_VALUE = 0,
  SDG_FIELD_BASIS_GRADIENT = 1,
  SDG_FIELD_BASIS_FACE_VALUE = 2,
  SDG_FIELD_BASIS_LSQ_INVERSE = 3,
  SDG_FIELD_QUAD_WEIGHT = 4,
  SDG_FIELD_FACE_QUAD_WEIGHT = 5,
  SDG_FIELD_QUAD_COORD = 6,
  SDG_FIELD_JACOBIAN_DET_W = 7,
  SDG_FIELD_JACOBIAN_TIW = 8,
  SDG_FIELD_MASS_INVERSE = 9,
  SDG_FIELD_MIN_EDGE = 10,
  SDG_FIELD_FACE_NORMAL = 11,
  SDG_FIELD_FACE_JACOBIAN_W = 12,
  SDG_FIELD_FACE_COORD = 13,
  SDG_FIELD_FACE_PARENT = 14,
  SDG_FIELD_FACE_LOCAL = 15,
  SDG_FIELD_FACE_SEQUENCE = 16,
  SDG_FIELD_FACE_ROTATION = 17,
  SDG_FIELD_QUAD_POINT = 18,      /* N_q x D reference-element volume points */
  SDG_FIELD_FACE_QUAD_POINT = 19, /* N_fq x (D-1) reference face points      */
  SDG_FIELD_BOUNDARY_FACE_COORD = 20,
  SDG_FIELD_BOUNDARY_FACE_NORMAL = 21,
  SDG_FIELD_BOUNDARY_FACE_JACOBIAN_W = 22,
  SDG_FIELD_BOUNDARY_FACE_PARENT = 23,
  SDG_FIELD_BOUNDARY_FACE_TYPE = 24,
  SDG_FIELD_ELEMENT_TYPE = 25,    /* (int32) E: enum sdg_element per element */
  SDG_FIELD_INNER_RADIUS = 26,
  SDG_FIELD_NODE_TAG = 27,
  SDG_FIELD_NODAL_VALUE = 28,
  SDG_FIELD_NODAL_FACE_VALUE = 29,
  SDG_FIELD_BOUNDARY_FACE_PHYSICAL = 30,
  SDG_FIELD_COUNT_ = 31
};

/* Split grid (MeshModelEnum::TriangleQuadrangle meshes of the reference): a periodic structured
 * n[0] x n[1] grid on [lo, hi] whose cells are quadrangles or, each with probability
 * split_probability (a seeded splitmix64 Bernoulli draw per cell, cells x fastest), two triangles
 * along the diagonal from (lo corner) to (hi corner) of the cell. A grid of one element type is an
 * ordinary mesh for sdg_create; a mixed one runs through the batched solver. For mixed grids the
 * exported QUAD_COORD / JACOBIAN_DET_W list every element's own quadrature points in element order
 * and ELEMENT_TYPE gives the element types; other fields are not available. */
typedef struct sdg_split_desc {
  int polynomial_order;   /* P = 1..4                                                     */
  int n[2];               /* cells per direction (>= 2)                                   */
  double lo[2];
  double hi[2];
  double split_probability;
  uint64_t seed;
} sdg_split_desc;

/* Mesh file written by gmsh (MSH 4.1 ASCII): affine elements of one type -- 3-node triangles or
 * 4-node parallelogram quadrangles in 2-D, 8-node parallelepiped hexahedra in 3-D -- with the
 * boundary given by the elements of one dimension lower in physical groups; boundary_count pairs
 * map a physical tag to an enum sdg_boundary type (addBoundaryCondition<T>(physical index),
 * SystemControl.cpp:76-80). Elements of negative orientation are mirrored; $Periodic links
 * (translations) pair the faces on their slave nodes with the faces on the master nodes; curved
 * (higher-order) and mixed element types are refused. */
typedef struct sdg_gmsh_desc {
  const char* path;
  int polynomial_order;   /* P = 1..5 (triangles 1..4)                                    */
  int boundary_count;
  const int* physical_index;
  const int* boundary;
} sdg_gmsh_desc;

typedef struct sdg_mesh sdg_mesh;
typedef struct sdg_ctx sdg_ctx;
typedef struct sdg_batch sdg_batch;

/* ---- library ---------------------------------------------------------------------------- */
/* ABI version compiled into the library. */
int sdg_abi_version(void);
/* Last error message of a context, or of the last failing context-less call when ctx is NULL. */
const char* sdg_last_error(const sdg_ctx* ctx);

/* ---- mesh and operands (host only; no device work) --------------------------------------- */
/* Replaces Mesh<SC>::initializeMesh + readMeshElement (ReadControl.cpp:286-335) for structured
 * periodic meshes: elements, faces with periodic pairing (Adjacency.cpp:185-328), Jacobians,
 * mass-matrix inverses, normals, basis and quadrature tables. */
int sdg_mesh_create_structured(const sdg_structured_desc* desc, sdg_mesh** out);
int sdg_mesh_create_split(const sdg_split_desc* desc, sdg_mesh** out);
/* Replaces Mesh<SC>::initializeMesh + readMeshElement (ReadControl.cpp:286-335) for a mesh file:
 * element and face construction (Adjacency.cpp:61-183), geometry, tables as above. */
int sdg_mesh_create_gmsh(const sdg_gmsh_desc* desc, sdg_mesh** out);
/* A mesh given as host arrays (borrowed for the call only): what the reference reads out of the gmsh
 * model that a setMesh(path, generate_mesh_function) callback builds (ReadControl.cpp:247-335) --
 * node coordinates, the elements' vertex lists, the boundary faces with their types and physical
 * tags, and translational periodic node pairs (Adjacency.cpp:185-328). Affine elements of one type;
 * faces, geometry and tables are built as for a mesh file. */
typedef struct sdg_array_mesh_desc {
  int element;                      /* SDG_ELEM_TRIANGLE, _QUADRANGLE, _TETRAHEDRON, _PYRAMID or    */
                                    /* _HEXAHEDRON                                                 */
  int polynomial_order;             /* P = 1..5 (triangles, tetrahedra and pyramids 1..4)           */
  int geometry_order;               /* order of the element map: 1, or the gmsh order of the node   */
                                    /* lists (1..5 in 2-D, 1..2 in 3-D); 0 means 1                  */
  int64_t node_count;
  const double* nodes;              /* node_count x 3 coordinates (z = 0 in 2-D)                   */
  int64_t element_count;
  const int64_t* elements;          /* element_count x geometry nodes (vertices first), 0-based,    */
                                    /* gmsh node order                                             */
  int64_t boundary_face_count;
  const int64_t* boundary_faces;    /* count x face vertices (2 in 2-D, 3 for tetrahedra, 4 for     */
                                    /* hexahedra and pyramids, a triangle padded with -1), any order */
  const int* boundary_type;         /* enum sdg_boundary per boundary face                          */
  const int* boundary_physical;     /* physical tag per boundary face (NULL: -1)                    */
  int periodic_link_count;          /* translational periodic pairs of boundaries                   */
  const int64_t* periodic_link_size; /* per link: its number of (slave, master) node pairs          */
  const int64_t* periodic_pairs;    /* all links' pairs, concatenated, 0-based nodes               */
} sdg_array_mesh_desc;
int sdg_mesh_create_from_arrays(const sdg_array_mesh_desc* desc, sdg_mesh** out);
void sdg_mesh_destroy(sdg_mesh* mesh);
int sdg_mesh_get_info(const sdg_mesh* mesh, sdg_mesh_info* info);
/* Number of scalar entries of an exported field (0 for an unknown field). */
int64_t sdg_mesh_field_size(const sdg_mesh* mesh, int field);
/* Copy an exported field into out (double* or int32_t* per the table above); `count` is the
 * capacity of out in entries and must be >= sdg_mesh_field_size. */
int sdg_mesh_export(const sdg_mesh* mesh, int field, void* out, int64_t count);

/* ---- solver context (device) ------------------------------------------------------------- */
/* Replaces Solver<SC> construction + the device-resident part of initializeSolver
 * (InitialCondition.cpp:151-186): uploads operands to HBM in the structure-of-arrays layout.
 * The mesh may be destroyed after this call. */
int sdg_create(const sdg_problem_desc* problem, const sdg_mesh* mesh, sdg_ctx** out);
void sdg_destroy(sdg_ctx* ctx);

/* Initial condition from primitive values at the volume quadrature points (E x N_q x (D+2),
 * primitive index fastest): conserved = f(primitive) (VariableConvertor.cpp:341-366), then the
 * unweighted least-squares projection C = U_q V (V^T V)^-1 (InitialCondition.cpp:85-104). */
int sdg_set_initial_primitive(sdg_ctx* ctx, const double* primitive_at_quadrature);
/* Set / get the modal coefficients (reference layout, E x N_b x NV with NV fastest). */
int sdg_set_state(sdg_ctx* ctx, const double* coefficients);
int sdg_get_state(sdg_ctx* ctx, double* coefficients);

/* CFL time step from the current state: min over elements and points of
 * CFL * minEdge / ((|u| + c) (P+1)^2) (TimeIntegration.cpp:104-131). */
int sdg_compute_dt(sdg_ctx* ctx, double cfl, double* dt);
/* No counterpart in the reference. For callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid allocation): sdg_compute_dt with no allocation of any
 * kind -- the per-element minima and their minimum live in a fixed device array of the library
 * (one per device, 16 MiB, held until the process ends) and only the result is copied back.
 * Meshes of more than 2^21 elements are refused (SDG_E_INVALID). Calls that overlap on one device
 * share that array and race. Same result as sdg_compute_dt bit for bit (per-element minima that
 * are NaN are skipped here). */
int sdg_compute_dt_noalloc(sdg_ctx* ctx, double cfl, double* dt);

/* One evaluation of the spatial residual R(U) of the current state, before M^-1: the passes
 * calculateGardientQuadrature .. calculateResidual of stepSolver (TimeIntegration.cpp:340-346).
 * residual: reference layout (E x N_b x NV). Leaves the state unchanged. */
int sdg_residual(sdg_ctx* ctx, double* residual);

/* Boundary data of a mesh with boundary faces: the primitive state (rho, u, v[, w], T) that the
 * user's BoundaryCondition::calculatePrimitiveFromCoordinate gives at every boundary face point,
 * F_b x N_fq x NV (points from SDG_FIELD_BOUNDARY_FACE_COORD). Converted like the reference's
 * boundary_dummy_variable_ (BoundaryCondition.cpp:29-50); call it before stepping, and again
 * between steps for time-varying boundary data. */
int sdg_set_boundary_primitive(sdg_ctx* ctx, const double* primitive_at_boundary_points);
/* NS only: the lifted total-gradient coefficients w of the current state, (NV*D) x N_b per
 * element column-major, row var*D+dim (variable_gradient_basis_function_coefficient_,
 * TimeIntegration.cpp:200-228). */
int sdg_get_gradient(sdg_ctx* ctx, double* gradient);
/* Artificial viscosity only: the node values of the last evaluation (node_artificial_viscosity_,
 * SolveControl.cpp:302), mesh nodes long. They are formed at the start of every step
 * (TimeIntegration.cpp:336-338) and of every sdg_residual call. */
int sdg_get_artificial_viscosity(sdg_ctx* ctx, double* node_values);

/* Advance nsteps explicit steps of size dt: stepSolver (TimeIntegration.cpp:327-350) repeated,
 * including the residual monitor calculateRelativeError (TimeIntegration.cpp:279-324).
 * relative_error (may be NULL): NV doubles, the monitor after the last step taken.
 * nan_flag (may be NULL): set to 1 if the monitor was all-NaN after some step, in which case
 * stepping stops after that step (SystemControl.cpp:185-191); steps_taken (may be NULL). */
int sdg_step(sdg_ctx* ctx, double dt, int nsteps, double* relative_error, int* nan_flag, int* steps_taken);

/* Device timing of the last sdg_step call in milliseconds (HIP events on the context stream). */
int sdg_last_step_ms(const sdg_ctx* ctx, double* ms);
/* Device synchronisation of the context stream. */
int sdg_synchronize(sdg_ctx* ctx);

/* Measurement hook (no counterpart in the reference): advance nsteps steps launching every
 * kernel eagerly (no graph) with HIP events around each launch on the context stream, and write
 * the average duration in milliseconds of each kernel class into avg_ms[0..n_out):
 *   [0] stage kernel 1 (Euler: the whole stage; Navier-Stokes: gradient lifting + volume flux)
 *   [1] Navier-Stokes face-flux kernel (Euler: 0)
 *   [2] Navier-Stokes residual + update + next traces (Euler: 0)
 *   [3] residual-monitor reduction
 * The state advances exactly as with sdg_step. */
#define SDG_KERNEL_CLASSES 4
int sdg_time_kernels(sdg_ctx* ctx, double dt, int nsteps, double* avg_ms, int n_out);

/* ---- raw state files (checkpoint / restart) ----------------------------------------------------
 * The reference's raw format (RawBinary.cpp:42-191): an 8-byte header holding
 * ZSTD_compressBound(payload bytes), then one zstd frame (level 1) of the payload:
 *   per element, in mesh order: the NV x N_b coefficients (column-major, as sdg_get_state), and
 *     for Navier-Stokes the (NV D) x N_b lifted-gradient coefficients (as sdg_get_gradient);
 *   per boundary face: the same blocks of its parent element (RawBinary.cpp:95-123);
 *   node_count doubles of nodal artificial viscosity: with shock capturing node_count must be the
 *     mesh's node count and the values are those of the last step; without it they are zero.
 * Replaces Solver::writeRawBinary (RawBinary.cpp:165-191) and RawBinaryCompress::write (:42-55).
 * Navier-Stokes BR2 meshes with boundary faces are refused (SDG_E_UNSUPPORTED): their per-face
 * gradient entries (volume part plus that face's lift) are not formed by the kernels. */
int sdg_raw_binary_write(sdg_ctx* ctx, const char* path, int64_t node_count);
/* Start from a raw file (InitialCondition.cpp:41-83): SDG_IC_LAST_STEP reads a file of this
 * discretisation; SDG_IC_SPECIFIC_FILE reads a file of order P-1 and sets the higher modes to zero
 * (the hierarchic basis of order P-1 is the first N_b(P-1) functions of order P). Gradient blocks
 * and the trailing data are skipped. */
int sdg_raw_binary_read(sdg_ctx* ctx, const char* path, int initial_condition);
/* The container alone (no device needed): compress `bytes` bytes into a raw file, and read one
 * back (payload NULL: only *bytes is set). */
int sdg_raw_file_write(const char* path, const void* payload, int64_t bytes);
int sdg_raw_file_read(const char* path, void* payload, int64_t capacity, int64_t* bytes);

/* ---- VTU output (host only; no device work) ---------------------------------------------------
 * ViewVariableEnum (Enum.cpp:210-233), same order. */
enum sdg_view_variable {
  SDG_VIEW_DENSITY = 0, SDG_VIEW_VELOCITY = 1, SDG_VIEW_TEMPERATURE = 2, SDG_VIEW_PRESSURE = 3,
  SDG_VIEW_SOUND_SPEED = 4, SDG_VIEW_MACH_NUMBER = 5, SDG_VIEW_ENTROPY = 6, SDG_VIEW_VORTICITY = 7,
  SDG_VIEW_HEAT_FLUX = 8, SDG_VIEW_ARTIFICIAL_VISCOSITY = 9, SDG_VIEW_VELOCITY_X = 10, SDG_VIEW_VELOCITY_Y = 11,
  SDG_VIEW_VELOCITY_Z = 12, SDG_VIEW_MACH_NUMBER_X = 13, SDG_VIEW_MACH_NUMBER_Y = 14, SDG_VIEW_MACH_NUMBER_Z = 15,
  SDG_VIEW_VORTICITY_X = 16, SDG_VIEW_VORTICITY_Y = 17, SDG_VIEW_VORTICITY_Z = 18, SDG_VIEW_HEAT_FLUX_X = 19,
  SDG_VIEW_HEAT_FLUX_Y = 20, SDG_VIEW_HEAT_FLUX_Z = 21
};
/* Replaces View::stepView for the element field (Paraview.cpp:154-313, ViewVariable::get
 * VariableConvertor.cpp:756-882): `coefficients` (reference layout, E x N_b x NV, e.g. from
 * sdg_get_state or a raw file) and, for Vorticity and HeatFlux, the lifted-gradient coefficients
 * `gradient` (E x N_b x NV D, as sdg_get_gradient; may be NULL otherwise) are evaluated at each
 * element's equispaced nodes of order P and written to `path` as an ASCII VTK XML unstructured grid
 * of one VTK Lagrange cell per element (curve 68, triangle 69, quadrilateral 70, hexahedron 72, nodes
 * in VTK's order), with the point data `variables[0..variable_count)` named as the reference names
 * them; in 2-D and 3-D Velocity, MachNumber and HeatFlux (the temperature gradient, as the
 * reference) are 3-vectors, Vorticity too in 3-D. ArtificialViscosity needs node_values (mesh nodes
 * long). Field data TMSTEP = step and TimeValue = time. */
int sdg_write_view(const sdg_mesh* mesh, const sdg_problem_desc* problem, const double* coefficients,
                   const double* gradient, const double* node_values, const int* variables, int variable_count, int step,
                   double time, const char* path);
/* sdg_write_view with Density, Velocity, Pressure, Temperature, MachNumber and, when node_values is not
 * NULL, ArtificialViscosity. */
int sdg_write_vtu(const sdg_mesh* mesh, const sdg_problem_desc* problem, const double* coefficients,
                  const double* node_values, int step, double time, const char* path);

/* ---- discrete adjoint (Navier-Stokes, periodic meshes: quadrangles P3-P5 and hexahedra P2-P4 in FP64,
 * hexahedra P2-P4 in mixed precision: the mixed run is recorded and recomputed as it ran, and the
 * FP64 residual is differentiated at its states)
 * The reference has no adjoint; these entry points transpose its residual (R of sdg_residual,
 * SpatialDiscrete.cpp:194-1116) and its RK stage update (TimeIntegration.cpp:181-198), each as the
 * transpose of the kernel family that runs the forward stages (sum-factorised on hexahedra by
 * default). All arrays are in the reference layout (E x N_b x NV, NV fastest). */

/* out = (dR/dU)^T lambda at the current state. */
int sdg_vjp_residual(sdg_ctx* ctx, const double* lambda, double* out);

/* Advance nsteps steps of size dt from the current state, like sdg_step, while recording
 * checkpoints for sdg_adjoint_backward: at most `snapshots` states (>= 1; the initial state is
 * one of them) at the positions of a binomial checkpointing schedule (Griewank's revolve bound:
 * with s snapshots, n steps are reversed with r repetitions where C(s+r, s) >= n). With
 * snapshots >= nsteps every step state is kept and nothing is recomputed twice. Fails with
 * SDG_E_STATE if the monitor turns NaN. */
int sdg_adjoint_forward(sdg_ctx* ctx, double dt, int nsteps, int snapshots);

/* Reverse sweep of the last sdg_adjoint_forward: given lambda_final = dJ/dU(t_N), writes
 * lambda_initial = dJ/dU(t_0). With every step recorded, the stage states (and their traces) of
 * the recorded run are used directly; otherwise they are recomputed from the checkpoints. Leaves
 * the context state at the recorded initial state. */
int sdg_adjoint_backward(sdg_ctx* ctx, const double* lambda_final, double* lambda_initial);

/* sdg_adjoint_backward with both adjoints in device memory of the context's device (hipMalloc'd,
 * or CUDA tensors of torch), in the same [E][N_b][NV] layout: no host copy and no PCIe transfer,
 * for callers whose objective gradient is formed on the device. Returns once lambda_initial is
 * written (stream synchronised). SDG_E_INVALID for host or other-device pointers. No counterpart
 * in the reference (it has no adjoint). */
int sdg_adjoint_backward_device(sdg_ctx* ctx, const double* lambda_final, double* lambda_initial);

/* ---- batched solver (no counterpart in the reference: it runs one problem per process) -------
 * Many independent 2-D compressible Euler problems on split grids (triangles and quadrangles,
 * orders 1..4), advanced together: every RK stage of the whole batch is one persistent launch,
 * steps are replayed from hipGraphs. Per problem the arithmetic is the single-problem engine's
 * (same kernels' formulas, reference layouts). All problems share the problem descriptor
 * (physics, flux, RK scheme, precision); each has its own mesh, order, state and time step. */
int sdg_batch_create(const sdg_problem_desc* problem, const sdg_mesh* const* meshes, int count, sdg_batch** out);
void sdg_batch_destroy(sdg_batch* batch);
const char* sdg_batch_last_error(const sdg_batch* batch);
/* primitive state at the mesh's QUAD_COORD points -> projected coefficients (InitialCondition.cpp:85-104) */
int sdg_batch_set_initial_primitive(sdg_batch* batch, int problem, const double* primitive);
/* coefficients per element in mesh order, N_b(element) x NV, NV fastest */
int sdg_batch_set_state(sdg_batch* batch, int problem, const double* coefficients);
int sdg_batch_get_state(sdg_batch* batch, int problem, double* coefficients);
/* per problem: min over its elements of the CFL time step (TimeIntegration.cpp:104-131) */
int sdg_batch_compute_dt(sdg_batch* batch, double cfl, double* dt);
/* nsteps RK steps, problem i with time step dt[i] */
int sdg_batch_step(sdg_batch* batch, const double* dt, int nsteps);
int sdg_batch_last_step_ms(const sdg_batch* batch, double* ms);
/* Autotuner: picks, per (element type, order) variant, the work-item size in element tiles and the
 * persistent grid size by timing stages on the current state (restored afterwards); stage_ms
 * (may be NULL) receives the best full-stage time. */
int sdg_batch_tune(sdg_batch* batch, double* stage_ms);
/* Current configuration: work-item tiles and elements per tile per variant (code = 4 * (quadrangle)
 * + P - 1; 0 where unused), and the grid size. */
int sdg_batch_get_config(const sdg_batch* batch, int* chunk_tiles, int* elements_per_tile, int* grid);

#ifdef __cplusplus
}
#endif

#endif /* SDG_H_ */
