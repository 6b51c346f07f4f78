// Command-line program of the front end, the counterpart of the reference's
// main (raytracer/main.cu:4-17): renderer2 <config.json> [output directory].
#include <cstdio>
#include "../../include/r2.h"

int main(int argc, char** argv) {
  if (argc < 2) {
    fprintf(stderr, "Error: Provide valid path to config file.\n");
    return 1;
  }
  r2_pipeline_report rep{};
  char err[1024] = {0};
  const int rc = r2_pipeline_run(argv[1], argc > 2 ? argv[2] : ".", 0, &rep, err, (int)sizeof(err));
  if (rc != R2_OK) {
    fprintf(stderr, "Error (%d): %s\n", rc, err);
    return 1;
  }
  if (rep.lods_filtered > 0) printf("Filtering took %.3f s (%d LODs)\n", rep.filter_ms * 1e-3, rep.lods_filtered);
  printf("Rendered %dx%d, %d spp, %d media in %.3f s (%.1f M paths/s)\n", rep.width, rep.height, rep.spp,
         rep.media_rendered, rep.render_ms * 1e-3, rep.mpaths_per_s);
  return 0;
}
