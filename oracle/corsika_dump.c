/* corsika_dump.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Reads a CORSIKA IACT (eventio) file with the reference's own reader
 * library (src/bernlohr/ of the reference tree, compiled from there by
 * oracle/build_ref.sh into oracle/_ref/) and prints what
 * ACorsikaIACTFile::Open and ::ReadEvent (src/ACorsikaIACTFile.cxx:190-470)
 * take from it: the telescope positions, and per event the wavelength bounds
 * (event header words 95 and 96) and every photon bunch with its telescope
 * and array numbers, cz computed as the reader computes it (:452).
 * Output lines:
 *   T <ntel> then one line per telescope: P <i> <x> <y> <z> <r>
 *   E <event number> <lambda_min> <lambda_max>
 *   B <array> <tel> <x> <y> <zem> <ctime> <cx> <cy> <cz> <lambda> <photons>
 * floats printed with %.9g (exact for float32). */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "initial.h"
#include "io_basic.h"
#include "mc_tel.h"
#include "fileopen.h"

#define MAX_TEL 1000
#define MAX_BUNCHES 1000000

int main(int argc, char** argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: corsika_dump FILE\n");
    return 2;
  }
  IO_BUFFER* iobuf = allocate_io_buffer(0);
  iobuf->max_length = 200000000;
  if ((iobuf->input_file = fileopen(argv[1], "r")) == NULL) {
    fprintf(stderr, "cannot open %s\n", argv[1]);
    return 1;
  }
  static double tx[MAX_TEL], ty[MAX_TEL], tz[MAX_TEL], tr[MAX_TEL];
  int ntel = 0;
  struct bunch* bunches = (struct bunch*)malloc(MAX_BUNCHES * sizeof(struct bunch));
  IO_ITEM_HEADER block;
  for (;;) {
    if (find_io_block(iobuf, &block) != 0) break;
    if (read_io_block(iobuf, &block) != 0) break;
    switch (block.type) {
      case IO_TYPE_MC_TELPOS:
        read_tel_pos(iobuf, MAX_TEL, &ntel, tx, ty, tz, tr);
        printf("T %d\n", ntel);
        for (int i = 0; i < ntel; i++) printf("P %d %.17g %.17g %.17g %.17g\n", i, tx[i], ty[i], tz[i], tr[i]);
        break;
      case IO_TYPE_MC_EVTH: {
        bernlohr_real evth[273];
        read_tel_block(iobuf, IO_TYPE_MC_EVTH, evth, 273);
        printf("E %.9g %.9g %.9g\n", (double)evth[1], (double)evth[95], (double)evth[96]);
        break;
      }
      case IO_TYPE_MC_TELARRAY: {
        IO_ITEM_HEADER ih;
        int array_no;
        begin_read_tel_array(iobuf, &ih, &array_no);
        for (int i = 0; i < ntel; i++) {
          IO_ITEM_HEADER sub;
          sub.type = IO_TYPE_MC_PHOTONS;
          if (search_sub_item(iobuf, &ih, &sub) < 0) break;
          int arr, tel, nb;
          double photons;
          if (read_tel_photons(iobuf, MAX_BUNCHES, &arr, &tel, &photons, bunches, &nb) < 0) continue;
          if (tel < 0) continue;
          for (int j = 0; j < nb; j++) {
            float cx = bunches[j].cx, cy = bunches[j].cy;
            float cz = -1. * sqrt(1. - cx * cx - cy * cy);
            printf("B %d %d %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g\n", arr, tel, bunches[j].x, bunches[j].y,
                   bunches[j].zem, bunches[j].ctime, cx, cy, cz, bunches[j].lambda, bunches[j].photons);
          }
        }
        end_read_tel_array(iobuf, &ih);
        break;
      }
      default:
        break;
    }
  }
  free(bunches);
  free_io_buffer(iobuf);
  return 0;
}
