/*
 * cac_io.h -- native readers and writers of the CAC on-disk formats (host
 * code in libcac_hip.so; no device and no context needed).
 *
 * - Data files: the reference's read_data format, read by ReadData::command
 *   (read_data.cpp:70-689) with Element::data_elements / data_nodes /
 *   set_nae / set_numint / set_intp (element.cpp:195-330, 478-540,
 *   1086-1166) and AtomVecAtomic::data_atom (atom_vec_atomic.cpp:49-71).
 *   Large sections are parsed in parallel chunks, numbers written with 17
 *   significant digits so a written and re-read system is the same bit for
 *   bit.
 * - Dumps: DumpAtom (dump_atom.cpp:35-157) and DumpElement
 *   (dump_element.cpp:72-263), one snapshot per call.
 *
 * Status codes are cac_hip.h's; cac_io_last_error() returns the message of
 * the calling thread's last failure (the reference's Error::all texts for
 * malformed data files).
 */
#ifndef CAC_IO_H
#define CAC_IO_H

#include <stdint.h>

#include "cac_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Message of this thread's last failed cac_data_* / cac_dump_* call. */
const char *cac_io_last_error(void);

/* A parsed data file.  Arrays are row-major; absent optional parts are
 * NULL.  Types are 1-based: entry t-1 of a per-type array belongs to type t.
 * masses[t-1] is NaN for a type without a Masses line; nae[3(t-1)..] is
 * {0,0,0} and numint[t-1] is -1 for an element type without NAAE / Numint
 * lines.  The Intp rows of all element types follow each other, types in
 * ascending order, each type's numint[t-1] rows in point order. */
typedef struct cac_data_view {
  const char *comment; /* line 1 */
  double lo[3], hi[3];
  int ntypes, entypes;
  const double *masses;   /* [ntypes] or NULL (no Masses section) */
  const int *nae;         /* [entypes][3] or NULL (no NAAE section) */
  const int *numint;      /* [entypes] or NULL (no Numint section) */
  int64_t nintp;          /* Intp rows */
  const int *intp_ijk;    /* [nintp][3] lattice indices, or NULL */
  const double *intp_w;   /* [nintp] weights, or NULL */
  int64_t natoms, nelements;
  const int64_t *atom_tag; /* [natoms] */
  const int *atom_type;    /* [natoms] */
  const double *x;         /* [natoms][3] */
  const int *atom_image;   /* [natoms][3] or NULL (5-word Atoms lines) */
  const int64_t *elem_tag; /* [nelements] */
  const int *elem_type;    /* [nelements] */
  const double *xc;        /* [nelements][3] element centers */
  const int *elem_image;   /* [nelements][3] or NULL */
  const double *nodex;     /* [nelements][8][3] */
  const double *nodev;     /* [nelements][8][3], zero where the Nodes lines carry no velocities */
  const int64_t *nodetag;  /* [nelements][8] */
  int node_velocities;     /* 1: the Nodes lines had 9 words (velocities) */
} cac_data_view;

typedef struct cac_data cac_data;

/* Read and parse a data file (replaces ReadData::command,
 * read_data.cpp:70-240: header :277-368, sections :155-215 and :379-437).
 * On success *out owns the parsed arrays until cac_data_free.  Errors:
 * CAC_ERR_IO (cannot open/read), CAC_ERR_DATA (malformed; the message is
 * the reference's, e.g. "Unexpected end of data file", "Unknown identifier
 * in data file: ...", "Incorrect atom format in data file"). */
int cac_data_read(const char *path, cac_data **out);
const cac_data_view *cac_data_get(const cac_data *d);
void cac_data_free(cac_data *d);

/* Write a data file in the reference's format: header, then Masses, NAAE,
 * Numint, Intp, Elements, Nodes (with node velocities if
 * node_velocities != 0; needs v->nodev) and Atoms, each section keyword
 * followed by one blank line (the line parse_keyword skips,
 * read_data.cpp:397).  Sections with no rows are left out. */
int cac_data_write(const char *path, const cac_data_view *v, int node_velocities);

/* One DumpAtom snapshot (dump_atom.cpp:35-157): the ITEM header (box
 * bounds "%g %g", boundary letters p or f per periodic[]), then one line
 * "id type x y z fx fy fz vx vy vz" (%d %d, then %g) per atom and one per
 * element node with type 1 (DumpAtom::epack_noscale_noimage).  The
 * NUMBER OF ATOMS item is the number of lines written: atoms + 8 elements
 * (the reference prints its centro-symmetry window count there, dump.cpp
 * :247, 353-366, while packing every atom and node).  A '*' in path is
 * replaced by the step (Dump::openfile, dump.cpp:190-206); append != 0 adds
 * the snapshot to an existing file.  tag NULL: 1..n; type NULL: 1;
 * nodetag NULL: 1..8 ne.  Arrays [n][3] / [ne][8][3]. */
int cac_dump_atom_write(const char *path, int append, int64_t step, const double lo[3], const double hi[3],
                        const int periodic[3], int64_t n, const int64_t *tag, const int *type, const double *x,
                        const double *f, const double *v, int64_t ne, const int64_t *nodetag, const double *nodex,
                        const double *nodef, const double *nodev);

/* One DumpElement snapshot (dump_element.cpp:72-263): the Tecplot header
 * (title, variables "x" .. "fz", zone t="load step N", n = 8 ne nodes,
 * e = ne elements, datapacking=point, zonetype=febrick), one line
 * "x y z vx vy vz fx fy fz" (%g) per node, elements in order, then one
 * connectivity line per element: the 1-based line numbers of its 8 nodes,
 * 8e+1 .. 8e+8.  The node order of the shape functions (element.cpp
 * :1341-1348: corners (-,-,-), (+,-,-), (+,+,-), (-,+,-), then the same at
 * +z) is the brick order such a zone expects.  The reference packs node
 * tags for that block (DumpElement::epack) but never writes it.
 * path, append as cac_dump_atom_write. */
int cac_dump_element_write(const char *path, int append, int64_t step, int64_t ne, const double *nodex,
                           const double *nodev, const double *nodef);

#ifdef __cplusplus
}
#endif

#endif /* CAC_IO_H */
