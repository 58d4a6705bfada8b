function [data,raw] = xensemblegpu(npar,input)
%   [data,raw] = XENSEMBLEGPU(npar,input) is xensemble (Simulator/xensemble.m)
%   for a project whose forms are named in p.gpu: all ensembles(2) x checks
%   paths of parallel ensemble npar run in one launch of the MI355X library
%   (xsh_mex('run_ensemble')); the outputs p.output{o}, the order
%   extrapolation and the D, D^2 sums of xensemble.m:84-105 stay here.
%   The library returns each path's averages as the p.d.av{o} arrays of
%   xpath.m:79-83 one after another in MATLAB order, so a reshape recovers
%   them.  If the library refuses the project (xSPDE:gpu:unsupported) this
%   falls back to the CPU xensemble.
%   Called by xsim (Simulator/xsim.m:53,59) in place of xensemble when
%   isfield(p,'gpu').  Not executed in the repository's tests (MATLAB is not
%   part of its image); the same bookkeeping in Python is
%   xspde_matlab_amd/xsim.py ensemble_moments, tested against the oracle.

if numel(input) > 1
  [data,raw] = xensemble(npar,input);            % sequences stay on the CPU
  return;
end
p = input{1};
try
  h = xsh_mex('create', xgpudesc(p));
catch err
  if strcmp(err.identifier,'xSPDE:gpu:unsupported')
    [data,raw] = xensemble(npar,input);
    return;
  end
  rethrow(err);
end
cleanup = onCleanup(@() xsh_mex('destroy', h));
info = xsh_mex('query', h);
xsh_mex('run_ensemble', h, npar);
if xsh_mex('nonfinite', h)                       % the kernels' divergence check
  warning('xSPDE:gpu:nonfinite', '%s: a field was not finite at a stored point (ensemble %d)', p.name, npar);
end
avall = xsh_mex('read_averages', h);             % [avlen, ensembles(2)*nchk]
raw = cell(1,p.nchk,p.ensembles(2));
if p.rawdata                                     % raw{s,nc,j} of xensemble.m:79
  rawall = xsh_mex('read_raw', h);               % [prod(p.d.raw{1}), ensembles(2)*nchk]
  for j = 1:p.ensembles(2)
    for nc = 1:p.nchk
      raw{1,nc,j} = {reshape(rawall(:,(j-1)*p.nchk+nc), p.d.raw{1})};
    end
  end
end
ne = 1./(p.ensembles(2)*p.ensembles(3)) - 1;     % ensemble scaling factor
m  = p.fbsw+p.nchk;
m1 = max(3,m+1);
data = cell(1);
for o = p.outputs
  data{1}{o} = zeros([p.d.data{o},m1]);
end
nobs = numel(p.gpu.observe);
for j = 1:p.ensembles(2)
  D = cell(p.nchk, max(p.outputs));
  for nc = 1:p.nchk
    col = avall(:,(j-1)*p.nchk+nc);
    av = cell(1,nobs);
    for o = 1:nobs
      n = info.points*info.obs_sites(o);
      av{o} = reshape(col(info.obs_offset(o)+(1:n)), p.d.av{o});
    end
    for o = p.outputs
      D{nc,o} = reshape(real(p.output{o}(av,p)),[p.d.data{o},1]);
    end
  end
  for o = p.outputs
    if p.checks(1) && p.order > 0                % extrapolating in time
      Del = (D{1,o} - D{2,o})/(2^(p.order)-1.);
      D{2,o} = D{1,o};
      D{1,o} = D{1,o}+Del;
    end
    nes = 1+ne;
    for n1 = 1:m
      data{1}{o}(:,n1) = data{1}{o}(:,n1) + D{n1,o}*nes;
    end
    data{1}{o}(:,m1) = data{1}{o}(:,m1)+D{1,o}.^2*nes;
  end
end
end
