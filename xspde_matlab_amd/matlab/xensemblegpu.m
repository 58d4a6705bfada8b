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
%   A project with p.fields > 1 runs one handle per field component
%   (xgpudesc(p,k)), all launched before any is read; observable o comes
%   from component p.gpu.ofield(o).
%   A sequence (xsim(p1,p2,...)) runs path by path: for every (j, nc) the
%   problems in order, each after the first from p.transfer of the previous
%   problem's final fields (xensemble.m:59-63; the randoms argument is empty on
%   this path), problem s on the random stream keyed by s - 1.
%   Called by xsim (Simulator/xsim.m:53,59) in place of xensemble when
%   isfield(p,'gpu').  Not executed in the repository's tests (MATLAB is not
%   part of its image); the same bookkeeping in Python is
%   xspde_matlab_amd/xsim.py ensemble_moments, tested against the oracle.

if numel(input) > 1
  [data,raw] = sequencegpu(npar,input);          % xsim(p1,p2,...): path by path
  return;
end
p = input{1};
F = p.fields{1};
if strcmp(p.gpu.deriv,'matrix'), F = 1; end      % coupled components: one descriptor
nobs = numel(p.gpu.observe);
of = ones(1,nobs);
if isfield(p.gpu,'ofield'), of = p.gpu.ofield; end
h = zeros(1,F,'uint64');
try
  for k = 1:F
    h(k) = xsh_mex('create', xgpudesc(p,k));
  end
catch err
  destroyall(h);                                 % the components created so far
  if strcmp(err.identifier,'xSPDE:gpu:unsupported')
    [data,raw] = xensemble(npar,input);
    return;
  end
  rethrow(err);
end
cleanup = onCleanup(@() destroyall(h));          % h holds every handle from here on
info = cell(1,F);
for k = 1:F
  info{k} = xsh_mex('query', h(k));
  xsh_mex('run_ensemble', h(k), npar);           % each on its own stream
end
avall = cell(1,F);
for k = 1:F
  if xsh_mex('nonfinite', h(k))                  % the kernels' divergence check
    warning('xSPDE:gpu:nonfinite', '%s: field %d was not finite at a stored point (ensemble %d)', ...
            p.name, k, npar);
  end
  avall{k} = xsh_mex('read_averages', h(k));     % [avlen, ensembles(2)*nchk]
end
raw = cell(1,p.nchk,p.ensembles(2));
if p.rawdata                                     % raw{s,nc,j} of xensemble.m:79
  sz = p.d.raw{1};
  rawall = cell(1,F);
  for k = 1:F
    rawall{k} = xsh_mex('read_raw', h(k));       % [prod(sz)/F, ensembles(2)*nchk]
  end
  for j = 1:p.ensembles(2)
    for nc = 1:p.nchk
      parts = cell(1,F);
      for k = 1:F
        parts{k} = reshape(rawall{k}(:,(j-1)*p.nchk+nc), [1, sz(2:end)]);
      end
      raw{1,nc,j} = {cat(1, parts{:})};
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
for j = 1:p.ensembles(2)
  D = cell(p.nchk, max(p.outputs));
  for nc = 1:p.nchk
    av = cell(1,nobs);
    for o = 1:nobs
      k = min(of(o), F);                           % coupled components: all in handle 1
      i = o;
      if F > 1, i = nnz(of(1:o) == k); end       % o's place among component k's observables
      col = avall{k}(:,(j-1)*p.nchk+nc);
      n = info{k}.points*info{k}.obs_sites(i);
      av{o} = reshape(col(info{k}.obs_offset(i)+(1:n)), p.d.av{o});
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

function [data,raw] = sequencegpu(npar,input)
% xensemble.m:25-108 over a sequence, each path on the GPU (xsh_mex('run_path'))
seq = numel(input);
p1 = input{1};
h = zeros(1,seq,'uint64');
try
  for s = 1:seq
    p = input{s};
    if p.fields{1} > 1 && ~strcmp(p.gpu.deriv,'matrix')
      error('xSPDE:gpu:unsupported','sequences of uncoupled field components stay on the CPU');
    end
    d = xgpudesc(p);
    d.seq = s-1;
    h(s) = xsh_mex('create', d);
  end
catch err
  destroyall(h);
  if strcmp(err.identifier,'xSPDE:gpu:unsupported')
    [data,raw] = xensemble(npar,input);
    return;
  end
  rethrow(err);
end
cleanup = onCleanup(@() destroyall(h));
info = cell(1,seq);
for s = 1:seq, info{s} = xsh_mex('query', h(s)); end
ne = 1./(p1.ensembles(2)*p1.ensembles(3)) - 1;   % ensemble scaling factor
m  = p1.fbsw+p1.nchk;
m1 = max(3,m+1);
data = cell(1,seq);
for s = 1:seq
  for o = input{s}.outputs
    data{s}{o} = zeros([input{s}.d.data{o},m1]);
  end
end
raw = cell(seq,p1.nchk,p1.ensembles(2));
for j = 1:p1.ensembles(2)
  D = cell(seq, p1.nchk);
  for nc = 1:p1.nchk
    b = [];
    for s = 1:seq
      p = input{s};
      a0 = [];
      if s > 1                                   % b: [nx*ny, E] -> p.d.ca of the previous problem
        bprev = reshape(b, input{s-1}.d.ca{1});
        a0 = complex(p.transfer{1}(bprev, [], p) + zeros(p.d.ca{1}));
      end
      [col, b] = xsh_mex('run_path', h(s), npar, j, nc, a0);
      nobs = numel(p.gpu.observe);
      av = cell(1,nobs);
      for o = 1:nobs
        n = info{s}.points*info{s}.obs_sites(o);
        av{o} = reshape(col(info{s}.obs_offset(o)+(1:n)), p.d.av{o});
      end
      for o = p.outputs
        D{s,nc}{o} = reshape(real(p.output{o}(av,p)),[p.d.data{o},1]);
      end
    end
  end
  for s = 1:seq
    p = input{s};
    for o = p.outputs
      if p.checks(1) && p.order > 0              % extrapolating in time
        Del = (D{s,1}{o} - D{s,2}{o})/(2^(p.order)-1.);
        D{s,2}{o} = D{s,1}{o};
        D{s,1}{o} = D{s,1}{o}+Del;
      end
      nes = 1+ne;
      for n1 = 1:m
        data{s}{o}(:,n1) = data{s}{o}(:,n1) + D{s,n1}{o}*nes;
      end
      data{s}{o}(:,m1) = data{s}{o}(:,m1)+D{s,1}{o}.^2*nes;
    end
  end
end
end

function destroyall(h)
for k = 1:numel(h)
  if h(k) ~= 0, xsh_mex('destroy', h(k)); end
end
end
