function H = ebSolveCyclesGpu(H)
%EBSOLVECYCLESGPU Run one solve stage of H on the MI355X library.
%   Same contract as EllipticBSAM/src/ebSolveCycles.m: on return every
%   level's Q holds the post-fill-down solution with ghosts filled, and one
%   stage record (amrLevels, nlev, vcycles, resHist, finalRel,
%   visibleCells, totalCells, avgFactor) is appended to H.stats.stages.
%   The record also carries gpuMs, the device time of the cycle loop.
%
%   H.opts.Precision = 'mixed' selects FP32 inner V-cycles inside FP64
%   defect correction (absent field: plain FP64).

h = bsamgpu_mex('create', H);
cleaner = onCleanup(@() bsamgpu_mex('destroy', h));
stage = bsamgpu_mex('solve', h, H.opts);
Q = bsamgpu_mex('get_q', h);
for k = 1:H.nlev
  for p = 1:H.lev{k}.np
    H.lev{k}.Q{p} = Q{k}{p};
  end
end
H.stats.stages{end+1} = stage;
if H.opts.Verbose >= 1
  fprintf(['[ebsam] stage %d: %d AMR level(s), %d V-cycle(s), ' ...
           '|r|/|f| = %.2e, avg factor %.3f, %d visible cells (gpu %.2f ms)\n'], ...
    numel(H.stats.stages), stage.amrLevels, stage.vcycles, stage.finalRel, ...
    stage.avgFactor, stage.visibleCells, stage.gpuMs);
end
end
