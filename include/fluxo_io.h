/*
 * fluxo_io: the reference's HDF5 file formats on the host (libfluxo_io.so).
 *
 *   HOPR mesh files (ReadMesh, src/mesh/mesh_readin.f90:168-669, and the side numbering of
 *   setLocalSideIDs / fillMeshInfo, src/mesh/prepare_mesh.f90:59-792) for conforming meshes
 *   whose boundaries are all periodic: the connectivity ElemToSide / SideToElem and the
 *   element node coordinates NodeCoords that fx_mesh_upload's metrics are computed from;
 *   state files (WriteState / WriteAnyState / GenerateFileSkeleton,
 *   src/io_hdf5/hdf5_output.f90:74-165, 559-614, 685-702, and Restart,
 *   src/restart/restart.f90:147-232, for the same N and node type).
 *
 * Layouts are the reference's (Fortran column-major, ids 1-based):
 *   ElemToSide(2,6,nElems), SideToElem(5,nSides), NodeCoords(3,0:Ngeo,0:Ngeo,0:Ngeo,nElems),
 *   U(nVar,0:N,0:N,0:N,nElems).
 * Every function returns 0 or a negative FXIO_E* code; fxio_last_error() has the text
 * (thread-local).  Host code only: nothing here touches the GPU.
 */
#ifndef FLUXO_IO_H
#define FLUXO_IO_H
#ifdef __cplusplus
extern "C" {
#endif

enum { FXIO_OK = 0, FXIO_EINVAL = -1, FXIO_EFILE = -2, FXIO_EFORMAT = -3, FXIO_EUNSUPPORTED = -4 };

const char *fxio_last_error(void);

/* Sizes of a HOPR mesh file: elements, unique sides (3 nElems when every side is an inner or
   periodic side), geometry degree, and 1 if its OutputNodeType is GAUSS-LOBATTO (0:
   EQUIDISTANT).  mesh_readin.f90:219-231, 448-459. */
int fxio_mesh_info(const char *file, int *nElems, int *nSides, int *Ngeo, int *nodeTypeGL);

/* ReadMesh + setLocalSideIDs + fillMeshInfo: element sides in CGNS order with flip 0 for the
   side's master (oriented) element and MOD(SIDE_Flip,10) for the other, periodic
   boundaries connected (BC_TYPE 1), side ids numbered in element/local-side order as the
   reference numbers inner sides.  Mortars, non-periodic boundaries and element types other
   than hexahedra are rejected (FXIO_EUNSUPPORTED). */
int fxio_mesh_read(const char *file, int *ElemToSide, int *SideToElem, double *NodeCoords);

/* State file of the reference's format: dataset DG_Solution (nVar,N+1,N+1,N+1,nElems) and the
   attributes Program, File_Type ("State"), Project_Name, File_Version, N, Time, MeshFile,
   NodeType ("GAUSS-LOBATTO"), VarNames, NComputation.  varNames: nVar strings. */
int fxio_state_write(const char *file, const char *projectName, const char *meshFile, int N, int nVar,
                     int nElems, double time, const char *const *varNames, const double *U);

/* N, nVar, nElems and Time of a state file. */
int fxio_state_info(const char *file, int *N, int *nVar, int *nElems, double *time);

/* Restart read of DG_Solution into U (same N and node type GAUSS-LOBATTO; other N or node
   types, which the reference interpolates, are FXIO_EUNSUPPORTED). */
int fxio_state_read(const char *file, int N, int nVar, int nElems, double *U);

#ifdef __cplusplus
}
#endif
#endif
