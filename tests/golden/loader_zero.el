# comment line
0 1
1 2

2 0
% another comment
3 3
