/*
 * fluxo_io: the reference's HDF5 file formats on the host (libfluxo_io.so).
 *
 *   HOPR mesh files (ReadMesh, src/mesh/mesh_readin.f90:168-669, and the side numbering of
 *   setLocalSideIDs / fillMeshInfo, src/mesh/prepare_mesh.f90:59-792) for meshes, conforming
 *   or with mortars, whose boundaries are all periodic: the connectivity ElemToSide /
 *   SideToElem, MortarType / MortarInfo and the
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

/* Sizes of a HOPR mesh file: elements, sides (big mortar sides included: nMortarSides + the
   inner sides; 3 nElems on a conforming mesh whose sides are all inner or periodic), geometry
   degree, and 1 if its OutputNodeType is GAUSS-LOBATTO (0: EQUIDISTANT).
   mesh_readin.f90:219-231, 448-459; prepare_mesh.f90:133-227. */
int fxio_mesh_info(const char *file, int *nElems, int *nSides, int *Ngeo, int *nodeTypeGL);

/* ReadMesh + setLocalSideIDs + fillMeshInfo (mesh_readin.f90:299-440, prepare_mesh.f90:133-227,
   662-726): element sides in CGNS order with flip 0 for the side's master (oriented) element
   and MOD(SIDE_Flip,10) for the other, periodic boundaries connected (BC_TYPE 1), side ids
   numbered as the reference numbers them: the big mortar sides first (1..nMortarSides), then
   the inner sides in element / local side / small-mortar-side order.  A big mortar side has
   its element as master and no slave; a small side of a mortar has the small element as
   slave and no master (SideToElem -1): its master data come from the big side.
   Non-periodic boundaries and element types other than hexahedra are rejected
   (FXIO_EUNSUPPORTED). */
int fxio_mesh_read(const char *file, int *ElemToSide, int *SideToElem, double *NodeCoords);

/* The non-conforming interfaces of the same numbering (fillMeshInfo, prepare_mesh.f90:700-718):
   nMortarSides, and when the arrays are given MortarType(2,nSides) ([type, index] of each big
   mortar side: type 1 = 1->4, 2 = 1->2 in eta, 3 = 1->2 in xi, index its column in
   MortarInfo; 0 for a small side, -1 otherwise) and MortarInfo(2,0:4,nMortarSides) ([SideID,
   flip] of the big side (0) and of its small sides (1..4), -1 where a 1->2 mortar has none).
   mesh_vars.f90 MortarType / MortarInfo. */
int fxio_mesh_mortars(const char *file, int *nMortarSides, int *MortarType, int *MortarInfo);

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
