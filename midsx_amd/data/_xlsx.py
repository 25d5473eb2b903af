import zipfile, re, sys
import xml.etree.ElementTree as ET
NS={'m':'http://schemas.openxmlformats.org/spreadsheetml/2006/main'}
def read(path):
    z=zipfile.ZipFile(path)
    ss=[]
    if 'xl/sharedStrings.xml' in z.namelist():
        r=ET.fromstring(z.read('xl/sharedStrings.xml'))
        for si in r.findall('m:si',NS):
            ss.append(''.join(t.text or '' for t in si.iter('{%s}t'%NS['m'])))
    wb=ET.fromstring(z.read('xl/workbook.xml'))
    sheets=[s.get('name') for s in wb.find('m:sheets',NS)]
    out={}
    for i,name in enumerate(sheets):
        fn='xl/worksheets/sheet%d.xml'%(i+1)
        if fn not in z.namelist(): continue
        r=ET.fromstring(z.read(fn))
        rows={}
        for c in r.iter('{%s}c'%NS['m']):
            ref=c.get('r'); t=c.get('t'); v=c.find('m:v',NS)
            if v is None: continue
            val=v.text
            if t=='s': val=ss[int(val)]
            col=re.match(r'[A-Z]+',ref).group(); row=int(ref[len(col):])
            rows.setdefault(row,{})[col]=val
        out[name]=rows
    return out
if __name__=='__main__':
    d=read(sys.argv[1])
    for name,rows in d.items():
        if len(sys.argv)>2 and name!=sys.argv[2]: continue
        print('=====',name)
        for r in sorted(rows)[:80]:
            print(r, rows[r])
