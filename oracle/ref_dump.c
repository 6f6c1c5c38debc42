/*
 * Test infrastructure only: a capture hook linked into the reference serial
 * solver when it is compiled (unmodified, from its own sources) into
 * oracle/_ref/ by oracle/Makefile.
 *
 * The reference's Iteration() (serial/Iterate.c:378) ends by handing the whole
 * cell array to WriteFinalResults().  The recipe compiles serial/Iterate.c with
 * -DWriteFinalResults=ref_dump_final so that this call lands here; the hook
 * writes the raw `struct CellProps` array (serial/include/ShellFunctions.h:14-37)
 * to Results/cells.bin and then calls the reference's own WriteFinalResults
 * (serial/FileIO.c:354) so every text output is unchanged.
 *
 * cells.bin layout: int32 n, int32 m, int32 sizeof(struct CellProps), then
 * n*m structs in the reference's storage order (idx = j*n + i).
 */
#include <stdio.h>
#include "ShellFunctions.h"
#include "FilesIO.h"

void ref_dump_final(char* OutputFile, struct CellProps* Cells, int* n, int* m,
                    int* postproc_prog)
{
  FILE* fp = fopen("Results/cells.bin", "wb");
  if (fp) {
    int hdr[3] = {*n, *m, (int)sizeof(struct CellProps)};
    fwrite(hdr, sizeof(int), 3, fp);
    fwrite(Cells, sizeof(struct CellProps), (size_t)(*n) * (size_t)(*m), fp);
    fclose(fp);
  }
  WriteFinalResults(OutputFile, Cells, n, m, postproc_prog);
}
